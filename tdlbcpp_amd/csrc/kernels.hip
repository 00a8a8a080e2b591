// gfx950 kernels of the 18-velocity Eggels-Somers ComputeUnit.
//
// Mapping: one thread per interior cell, k (the fastest index of both layouts)
// on the lanes, one block per run of k in one (i,j) row.  In LIJK every slot
// load and store of a wavefront is a contiguous 256 B run; every memory
// location is read once and written once per step, so the kernels move the
// algorithmic 2 x 18 x sizeof(T) bytes per cell (+ Nu with LES).
//
// Reference semantics (file:line in the reference tree):
//   collision  CollisionEgglesSomers.hpp:26-181   moments  :185-256
//   Esotwist accessors  StreamingEsoteric.hpp:27-180
//   Simple push/pull    StreamingSimple.hpp:18-126   bounce-back Boundary.hpp:13-250
//   forcing             Forcing.hpp:56-176
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "codec16.h"
#include "device.h"
#include "es_collide.h"
#include "kernels.h"
#include "rushton.h"

namespace tdlb {

namespace {

constexpr int kBlock = 256;
// Stream-collide kernels take one block per interior k-row when the row fits
// (up to kRowBlock lanes).  The population arrays keep interior cells only in
// their rows (see device.h), so a block's run of each slot plane is whole
// lines that no other block touches; at 512^3 one 512-lane block per row
// measured 4 % faster than two 256-lane blocks in FP32 and 3-4 % in 16 bits.
constexpr int kRowBlock = 512;
template <typename S>
constexpr int row_block_max() { return kRowBlock; }

inline uint32_t blocks_for(uint64_t n, int bs = kBlock) {
    uint64_t b = (n + bs - 1) / bs;
    return (uint32_t)(b > 0x7fffffffull ? 0x7fffffffull : (b ? b : 1));
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b to XCD b % 8),
// each XCD with its own L2.  xcd_block renumbers blocks so that each XCD takes
// one contiguous range of rows.  In IJKL a cell's 18 slots share lines with its
// neighbours', which neighbouring rows then find in the same L2 (measured 13-20 %
// faster); in LIJK every line is read by one row only and the default order,
// all XCDs streaming through one region of HBM together, is faster (5 %).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
    constexpr uint32_t nx = kXcds;
    const uint32_t x = b % nx, q = b / nx;
    const uint32_t per = n / nx, rem = n % nx;
    return x * per + min(x, rem) + q;
}
template <int L>
__device__ __forceinline__ uint32_t layout_block() {
    return L == LayoutIJKL ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
}

// ------------------------------------------------------------------ precisions
// S: stored population type, T: arithmetic and F type, N: stored Nu type, D:
// type the reference's double literals promote to.  kShift: populations are
// stored as scaled deviations from the rest populations (see shift_scale) and
// Nu in the same 16-bit format (enc_nu / dec_nu).
struct PrecF32 {
    using S = float; using T = float; using N = float; using D = double;
    static constexpr bool kShift = false;
};
struct PrecF64 {
    using S = double; using T = double; using N = double; using D = double;
    static constexpr bool kShift = false;
};
// The 16-bit formats: enc rounds the scaled deviation s to the stored type
// (round to nearest even), dec widens it exactly; enc_dev / dec_dev take the
// unscaled deviation s / k of direction d (k = shift_scale(d), a power of two),
// also exactly.  Nu (>= 0, 1e-6 .. 1e-2 in a stirred tank) is stored in 16
// bits, scaled by 2^10 in fp16 so that it stays above fp16's smallest normal
// (6.1e-5) down to 6e-8.
struct PrecF16 {
    using S = _Float16; using T = float; using N = _Float16; using D = float;
    static constexpr bool kShift = true;
    __device__ static S enc(float s) { return (S)s; }
    __device__ static float dec(S x) { return (float)x; }
    __device__ static S enc_dev(float v, int d) { return enc(v * shift_scale(d)); }
    __device__ static float dec_dev(S x, int d) { return dec(x) * (1.0f / shift_scale(d)); }
    __host__ __device__ static N enc_nu(float v) { return (N)(v * 1024.0f); }
    __host__ __device__ static float dec_nu(N x) { return (float)x * (1.0f / 1024.0f); }
};
struct PrecBF16 {
    using S = __bf16; using T = float; using N = __bf16; using D = float;
    static constexpr bool kShift = true;
    __device__ static S enc(float s) { return (S)s; }
    __device__ static float dec(S x) { return (float)x; }
    __device__ static S enc_dev(float v, int d) { return enc(v * shift_scale(d)); }
    __device__ static float dec_dev(S x, int d) { return dec(x) * (1.0f / shift_scale(d)); }
    __host__ __device__ static N enc_nu(float v) { return (N)v; }
    __host__ __device__ static float dec_nu(N x) { return (float)x; }
};
// E3M12 (codec16.h): 12 fraction bits for the deviations, the scale k folded
// into the code's exponent; Nu as in PrecF16.
struct PrecE3M12 {
    using S = int16_t; using T = float; using N = _Float16; using D = float;
    static constexpr bool kShift = true;
    __device__ static S enc(float s) { return e3m12_encode<0>(s); }
    __device__ static float dec(S x) { return e3m12_decode<0>(x); }
    __device__ static S enc_dev(float v, int d) { return d < 6 ? e3m12_encode<1>(v) : e3m12_encode<2>(v); }
    __device__ static float dec_dev(S x, int d) { return d < 6 ? e3m12_decode<1>(x) : e3m12_decode<2>(x); }
    __host__ __device__ static N enc_nu(float v) { return PrecF16::enc_nu(v); }
    __host__ __device__ static float dec_nu(N x) { return PrecF16::dec_nu(x); }
};
// bit patterns moved unchanged
template <typename W>
struct RawBits {
    using S = W; using T = W; using N = W; using D = W;
    static constexpr bool kShift = false;
};

// Cache hints of the stream-collide accesses.  In LIJK a wavefront of the fused
// Esotwist kernel reads and writes most slot planes at its own k positions,
// lines that no other wavefront touches: those accesses carry the
// non-temporal hint (nt).  The planes it reads at k+1 share their edge line
// with the next wavefront, which finds it in the cache only without the hint.
// This holds only when every k-row of a slot plane is a whole number of lines:
// otherwise a row's first and last lines are shared with the neighbouring rows'
// blocks and the hint costs a refetch (line_rows).  Measured (Esotwist, LIJK):
// 512^3 FP32 4.4 % faster with the hint on the unshifted planes (on loads only
// 1.8 % slower, on every access 6 % slower), 320^3 6 % faster in FP32 and FP16;
// rows off the line grid 9-27 % slower (200^3, and 160^3 / 96^3 in 16 bits).
// IJKL (a line holds slots of several cells) is 35 % slower with it; the Simple
// pull shows no difference; batches, whose tiles mix tanks, take no hint.  The
// hint is a template argument: each instantiation holds one kind of access, and
// call sites pick the instantiation per plane (a runtime flag lets the compiler
// merge the two accesses of one address and drop the hint).
template <bool NT, typename S>
__device__ __forceinline__ S load_hint(const S* __restrict__ p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename S>
__device__ __forceinline__ void store_hint(S* __restrict__ p, S v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <typename PR, bool NT = false>
__device__ __forceinline__ typename PR::T load_pop(const typename PR::S* __restrict__ Q, uint64_t a, int d,
                                                   float rho0) {
    // (dec + c0) / k as one multiply-add: x / k and c0 / k are exact (k a power of two)
    if constexpr (PR::kShift)
        return fmaf(PR::dec(load_hint<NT>(Q + a)), 1.0f / shift_scale(d),
                    shift_offset(rho0) * (1.0f / shift_scale(d)));
    else return load_hint<NT>(Q + a);
}
template <typename PR>
__device__ __forceinline__ typename PR::T load_nu(const typename PR::N* __restrict__ Nu, uint64_t n) {
    if constexpr (PR::kShift) return PR::dec_nu(Nu[n]);
    else return Nu[n];
}
template <typename PR>
__device__ __forceinline__ void store_nu(typename PR::N* __restrict__ Nu, uint64_t n, typename PR::T v) {
    if constexpr (PR::kShift) Nu[n] = PR::enc_nu(v);
    else Nu[n] = v;
}
template <typename PR, bool NT = false>
__device__ __forceinline__ void store_pop(typename PR::S* __restrict__ Q, uint64_t a, int d, typename PR::T v,
                                          float rho0) {
    // k v - c0 as one multiply-add (k v is exact)
    if constexpr (PR::kShift) store_hint<NT>(Q + a, PR::enc(fmaf(v, shift_scale(d), -shift_offset(rho0))));
    else store_hint<NT>(Q + a, v);
}
// k-rows of the population planes start and end on 128-byte lines
template <typename S>
inline bool line_rows(const Grid& g) {
    return ((uint64_t)g.qz * sizeof(S)) % 128 == 0 && ((uint64_t)g.qsize * sizeof(S)) % 128 == 0 &&
           ((uint64_t)g.qyz * sizeof(S)) % 128 == 0;
}

// The fused collision of the 16-bit formats works on deviations from the rest
// populations c0 / k (es_collide_dev): f - c0 / k = dec(s) / k and the stored
// value of a deviation p' is k p', both exact (k a power of two), so the
// 16-bit rounding is the only one between collisions.
template <typename PR, bool NT = false>
__device__ __forceinline__ float load_dev(const typename PR::S* __restrict__ Q, uint64_t a, int d) {
    return PR::dec_dev(load_hint<NT>(Q + a), d);
}
template <typename PR, bool NT = false>
__device__ __forceinline__ void store_dev(typename PR::S* __restrict__ Q, uint64_t a, int d, float v) {
    store_hint<NT>(Q + a, PR::enc_dev(v, d));
}
// density of the rest populations c0 / k summed over the 18 directions
__device__ __forceinline__ float rest_density(float rho0) { return 6.0f * shift_offset(rho0); }
// the hardware's reciprocal and square root (1 ulp) for the float collision of
// the 16-bit formats; the IEEE-exact forms expand to ~9 instructions each
struct HwOps {
    __device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    __device__ static float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
};

// ------------------------------------------------------------------ fills
// every slot of the population array, row padding included
template <typename T, int L>
__global__ void k_fill_q(T* __restrict__ Q, Grid g, T rho) {
    const uint64_t cells = L == LayoutIJKL ? g.size : g.qsize;
    const uint64_t total = cells * (uint64_t)kQ;
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < total;
         a += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t slot = (L == LayoutIJKL) ? a % kQ : a / cells;
        Q[a] = slot == 0 ? rho : T(0);
    }
}

// Between the population array of the kernels (layout DL, padded rows) and the
// dense array of the boundary (layout BL: IJKL n*18 + l, LIJK l*size + n); one
// boundary element per thread.
template <typename W, int DL, int BL, bool TO_BOUNDARY>
__global__ void k_q_boundary(const W* __restrict__ src, W* __restrict__ dst, Grid g) {
    const uint64_t total = g.size * (uint64_t)kQ;
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < total;
         a += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t n, l;
        if (BL == LayoutLIJK) {
            l = a / g.size;
            n = a - l * g.size;
        } else {
            n = a / kQ;
            l = a - n * kQ;
        }
        int i, j, k;
        decode_cell(g, n, i, j, k);
        const uint64_t q = qaddr<DL>(g, qcell(g, i, j, k), (int)l);
        if (TO_BOUNDARY) dst[a] = src[q];
        else dst[q] = src[a];
    }
}

__global__ void k_zero_words(uint32_t* __restrict__ p, uint64_t n) {
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n; a += (uint64_t)gridDim.x * blockDim.x)
        p[a] = 0u;
}
__global__ void k_zero_bytes(uint8_t* __restrict__ p, uint64_t n) {
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n; a += (uint64_t)gridDim.x * blockDim.x)
        p[a] = 0;
}
__global__ void k_copy_words(uint32_t* __restrict__ d, const uint32_t* __restrict__ s, uint64_t n) {
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n; a += (uint64_t)gridDim.x * blockDim.x)
        d[a] = s[a];
}
__global__ void k_copy_bytes(uint8_t* __restrict__ d, const uint8_t* __restrict__ s, uint64_t n) {
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n; a += (uint64_t)gridDim.x * blockDim.x)
        d[a] = s[a];
}

// ------------------------------------------------------------------ Esotwist
// Direction d of cell c is stored at home(c,d) = c + (1 where e_d < 0, periodic
// on the interior), in slot d at even parity and opp(d) at odd parity.  read
// takes the 18 values from there; write puts population d where the read of
// opp(d) looks (StreamingEsoteric.hpp:82-130); writeMoments puts value d back
// where d was read.
// mout receives the moments of MODE_MOM: Q itself (in place) when populations
// are stored as T, a separate T array in the reference layout otherwise.
template <typename PR, int L, int MODE, int EVEN, bool HINT = false>
__device__ __forceinline__ void esotwist_cell(typename PR::S* __restrict__ Q, typename PR::T* mout,
                                              const typename PR::T* __restrict__ F, typename PR::N* __restrict__ Nu,
                                              const uint64_t* __restrict__ fmask, const Grid& g,
                                              const CollideParams<typename PR::T>& P, float rho0, int i, int j,
                                              int k) {
    using T = typename PR::T;
    using D = typename PR::D;
    const uint32_t n = cell_index(g, i, j, k);
    const uint32_t qn = qcell(g, i, j, k);
    const int32_t di = (i == g.xg1) ? -(int32_t)((uint32_t)(g.xg1 - 1) * g.qyz) : (int32_t)g.qyz;
    const int32_t dj = (j == g.yg1) ? -(int32_t)((uint32_t)(g.yg1 - 1) * g.qz) : (int32_t)g.qz;
    const int32_t dk = (k == g.zg1) ? -(g.zg1 - 1) : 1;

    uint32_t home[kQ];
#pragma unroll
    for (int d = 0; d < kQ; d++)
        home[d] = qn + (uint32_t)((twist_dx(d) ? di : 0) + (twist_dy(d) ? dj : 0) + (twist_dz(d) ? dk : 0));

    T v[kQ];
#pragma unroll
    for (int d = 0; d < kQ; d++) {
        const uint64_t a = qaddr<L>(g, home[d], EVEN ? d : opposite(d));
        // reduced-precision storage keeps the moments of MODE_COLL in mout
        // with HINT, the planes read at the cell's own k take the streaming hint
        if constexpr (PR::kShift && MODE == MODE_COLL) v[d] = mout[a];
        else if constexpr (MODE == MODE_FUSED && PR::kShift) {
            if (!HINT || twist_dz(d)) v[d] = load_dev<PR>(Q, a, d);
            else v[d] = load_dev<PR, true>(Q, a, d);
        } else if constexpr (MODE == MODE_FUSED) {
            if (!HINT || twist_dz(d)) v[d] = load_pop<PR>(Q, a, d, rho0);
            else v[d] = load_pop<PR, true>(Q, a, d, rho0);
        }
        else v[d] = load_pop<PR>(Q, a, d, rho0);
    }

    if (MODE == MODE_MOM) {
        T m[kQ];
        es_moments<T, D>(v, m);
#pragma unroll
        for (int d = 0; d < kQ; d++) mout[qaddr<L>(g, home[d], EVEN ? d : opposite(d))] = m[d];
        return;
    }

    T fx = T(0), fy = T(0), fz = T(0);
    if (mask_test(fmask, g, i, j, k)) {
        fx = F[3 * (uint64_t)n + 0];
        fy = F[3 * (uint64_t)n + 1];
        fz = F[3 * (uint64_t)n + 2];
    }
    T nuc = P.use_les ? load_nu<PR>(Nu, n) : T(0);
    T p[kQ];
    if constexpr (MODE == MODE_FUSED && PR::kShift) {
        es_collide_dev<HwOps>(v, rest_density(rho0), fx, fy, fz, nuc, P, p);
    } else if (MODE == MODE_FUSED) {
        T m[kQ];
        es_moments<T, D>(v, m);
        es_collide<T, D>(m, fx, fy, fz, nuc, P, p);
    } else {
        es_collide<T, D>(v, fx, fy, fz, nuc, P, p);
    }
    if (P.use_les) store_nu<PR>(Nu, n, nuc);
    // Recompute the 18 home cells instead of keeping them live through the
    // collision: the opaque copy stops the compiler from reusing the load
    // addresses, which frees ~36 VGPRs and raises occupancy.
    uint32_t n2 = qn;
    int32_t di2 = di, dj2 = dj, dk2 = dk;
    asm volatile("" : "+v"(n2), "+v"(di2), "+v"(dj2), "+v"(dk2));
#pragma unroll
    for (int d = 0; d < kQ; d++) {
        const int e = opposite(d);
        const uint32_t h = n2 + (uint32_t)((twist_dx(e) ? di2 : 0) + (twist_dy(e) ? dj2 : 0) + (twist_dz(e) ? dk2 : 0));
        const uint64_t a = qaddr<L>(g, h, EVEN ? e : opposite(e));
        if constexpr (MODE == MODE_FUSED) {
            if (!HINT || twist_dz(e)) {
                if constexpr (PR::kShift) store_dev<PR>(Q, a, d, p[d]);
                else store_pop<PR>(Q, a, d, p[d], rho0);
            } else {
                if constexpr (PR::kShift) store_dev<PR, true>(Q, a, d, p[d]);
                else store_pop<PR, true>(Q, a, d, p[d], rho0);
            }
        }
        else store_pop<PR>(Q, a, d, p[d], rho0);
    }
}

template <typename PR, int L, int MODE, int EVEN, bool HINT>
__global__ __launch_bounds__(kRowBlock) void k_esotwist(typename PR::S* __restrict__ Q, typename PR::T* mout,
                                                     const typename PR::T* __restrict__ F,
                                                     typename PR::N* __restrict__ Nu,
                                                     const uint64_t* __restrict__ fmask, Grid g,
                                                     CollideParams<typename PR::T> P, uint32_t nkb, float rho0) {
    const RowBlock rb = row_block(g, layout_block<L>(), nkb, blockDim.x);
    const int k = rb.k0 + (int)threadIdx.x;
    if (k > g.zg1) return;
    esotwist_cell<PR, L, MODE, EVEN, HINT>(Q, mout, F, Nu, fmask, g, P, rho0, rb.i, rb.j, k);
}

// ------------------------------------------------------------------ Esotwist, IJKL through LDS
// The fused step in IJKL, where a cell's 18 slots are one 72-byte record and a
// lane reading single slots spans a line per lane.  A workgroup owns a tile
// of RI x KT cells in (i, k) and marches along j through a chunk of planes,
// holding two slabs in LDS: the records of rows i0 .. i0 + RI and cells
// k0 .. k0 + KT (the extra row and cell are the tile's +1 halo, or the
// periodic wrap) of plane j and of plane j + 1.  A slab is read from HBM with
// consecutive lanes on consecutive words (whole records, coalesced), the cells
// of plane j take their 18 values from the two slabs and put the collided ones
// back in place, and the slab of plane j is written back, every word whose
// owner is a cell of this tile that has run: a slot of a record belongs to one
// cell (home(c, d) = c + twist(d), so the owner is the record - twist), and the
// words owned by another tile, or by a plane outside the chunk, are neither
// used nor written.  Two tiles may write disjoint words of one line; no word
// has two writers.
constexpr int kSlabPad = kQ + 1;  // LDS words per record: consecutive records start in different banks

template <int EVEN>
__device__ __forceinline__ uint32_t slot_twist_mask(int axis) {
    uint32_t m = 0;
#pragma unroll
    for (int s = 0; s < kQ; s++) {
        const int d = EVEN ? s : opposite(s);
        const int t = axis == 0 ? twist_dx(d) : axis == 1 ? twist_dy(d) : twist_dz(d);
        m |= (uint32_t)t << s;
    }
    return m;
}

template <typename PR, int EVEN, int RI, int KT>
__global__ __launch_bounds__(RI * KT) void k_esotwist_ijkl_lds(typename PR::S* __restrict__ Q,
                                                            const typename PR::T* __restrict__ F,
                                                            typename PR::N* __restrict__ Nu,
                                                            const uint64_t* __restrict__ fmask, Grid g,
                                                            CollideParams<typename PR::T> P, int nti, int ntk,
                                                            int jchunk, float rho0) {
    using T = typename PR::T;
    using D = typename PR::D;
    using S = typename PR::S;
    static_assert(!PR::kShift, "whole populations only");
    constexpr int NR = RI + 1, NK = KT + 1, REC = NR * NK, NT = RI * KT;
    __shared__ S slab[2][REC * kSlabPad];

    const uint32_t b = blockIdx.x;
    const int tk = (int)(b % (uint32_t)ntk);
    const int ti = (int)((b / (uint32_t)ntk) % (uint32_t)nti);
    const int tj = (int)(b / ((uint32_t)ntk * (uint32_t)nti));
    const int i0 = 1 + ti * RI, k0 = 1 + tk * KT;
    const int ri = min(RI, g.xg1 - i0 + 1), rk = min(KT, g.zg1 - k0 + 1);  // cells of the tile
    const int js = 1 + tj * jchunk;
    if (js > g.yg1) return;
    const int je = min(g.yg1, js + jchunk - 1);
    const int tid = (int)threadIdx.x;
    const uint32_t txm = slot_twist_mask<EVEN>(0), tym = slot_twist_mask<EVEN>(1), tzm = slot_twist_mask<EVEN>(2);

    // A slab row (records k0 .. k0 + rk of row i in plane j) is RW consecutive
    // words in HBM, except that the +1 record past the last interior cell is
    // the wrapped k = 1.  Lane t copies words c * NT + t of every row: its word
    // offsets, LDS positions and slot bits are fixed for the tile, and only the
    // row's base address changes with the row and the plane.
    constexpr int RW = NK * kQ, RC = (RW + NT - 1) / NT, ITER = NR * RC;
    int goff[RC], loff[RC];
    uint32_t wbits[RC];  // bit 0: word in the tile's slab, 1: twist in x, 2: in y, 3: owner k inside the tile
    const bool kwrap = k0 + rk > g.zg1;
#pragma unroll
    for (int c = 0; c < RC; c++) {
        const int wr = c * NT + tid, kk = wr / kQ, sl = wr - kk * kQ;
        const bool ok = wr < RW && kk <= rk;
        goff[c] = (kwrap && kk == rk) ? (1 - k0) * kQ + sl : wr;  // the wrapped record sits at k = 1
        loff[c] = kk * kSlabPad + sl;
        const int okk = kk - (int)((tzm >> sl) & 1u);
        wbits[c] = (ok ? 1u : 0u) | (((txm >> sl) & 1u) << 1) | (((tym >> sl) & 1u) << 2) |
                   ((okk >= 0 && okk < rk) ? 8u : 0u);
    }
    // word offset of record (i0 + r, j, k0): the row past the tile wraps to i = 1
    auto row_base = [&](int r, int j) __attribute__((always_inline)) -> int64_t {
        int i = i0 + r;
        if (i > g.xg1) i -= g.xg1;
        const int jp = j > g.yg1 ? j - g.yg1 : j;
        return (int64_t)cell_index(g, i, jp, k0) * kQ;
    };
    // every load of a slab issued before any is used
    auto fetch = [&](S (&reg)[ITER], int j) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (r > ri) break;
            const S* row = Q + row_base(r, j);
#pragma unroll
            for (int c = 0; c < RC; c++)
                if (wbits[c] & 1u) reg[r * RC + c] = row[goff[c]];
        }
    };
    auto put = [&](const S (&reg)[ITER], int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (r > ri) break;
#pragma unroll
            for (int c = 0; c < RC; c++)
                if (wbits[c] & 1u) slab[buf][r * NK * kSlabPad + loff[c]] = reg[r * RC + c];
        }
    };
    // write back the words of slab buf (plane j) owned by cells of this tile in
    // planes js .. last: owner row r - tx inside the tile, owner k inside the
    // tile, owner plane j - ty inside js .. last
    auto store = [&](int buf, int j, int last) __attribute__((always_inline)) {
        const bool pl0 = j >= js && j <= last, pl1 = j - 1 >= js && j - 1 <= last;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (r > ri) break;
            S* row = Q + row_base(r, j);
#pragma unroll
            for (int c = 0; c < RC; c++) {
                const uint32_t wb = wbits[c];
                const bool tx = wb & 2u, ty = wb & 4u;
                const bool rok = r == ri ? tx : (r == 0 ? !tx : true);
                if ((wb & 1u) && (wb & 8u) && rok && (ty ? pl1 : pl0))
                    row[goff[c]] = slab[buf][r * NK * kSlabPad + loff[c]];
            }
        }
    };

    const int r = tid / KT, kk = tid - (tid / KT) * KT;
    const bool active = r < ri && kk < rk;
    const int ci = i0 + (active ? r : 0), ck = k0 + (active ? kk : 0);
    // Nu and the F-occupancy word of a cell, fetched a plane ahead
    auto cell_pre = [&](int j, T& nu, uint64_t& mw) __attribute__((always_inline)) {
        nu = P.use_les ? load_nu<PR>(Nu, cell_index(g, ci, j, ck)) : T(0);
        mw = fmask[mask_word_index(g, ci, j, ck)];
    };
    auto collide = [&](int j, int A, int B, T nuc, uint64_t mw) __attribute__((always_inline)) {
        if (!active) return;
        const uint32_t n = cell_index(g, ci, j, ck);
        T v[kQ];
#pragma unroll
        for (int d = 0; d < kQ; d++) {
            const int at = ((r + twist_dx(d)) * NK + kk + twist_dz(d)) * kSlabPad + (EVEN ? d : opposite(d));
            v[d] = twist_dy(d) ? slab[B][at] : slab[A][at];
        }
        T fx = T(0), fy = T(0), fz = T(0);
        if ((mw >> (ck & 63)) & 1ull) {
            fx = F[3 * (uint64_t)n + 0];
            fy = F[3 * (uint64_t)n + 1];
            fz = F[3 * (uint64_t)n + 2];
        }
        T m[kQ], p[kQ];
        es_moments<T, D>(v, m);
        es_collide<T, D>(m, fx, fy, fz, nuc, P, p);
        if (P.use_les) store_nu<PR>(Nu, n, nuc);
#pragma unroll
        for (int d = 0; d < kQ; d++) {
            const int e = opposite(d);
            const int at = ((r + twist_dx(e)) * NK + kk + twist_dz(e)) * kSlabPad + (EVEN ? e : opposite(e));
            if (twist_dy(e)) slab[B][at] = p[d];
            else slab[A][at] = p[d];
        }
    };
    // Two slabs in LDS (planes j, j + 1) and one in flight in registers: plane
    // j + 2 is requested while plane j collides and fills the buffer of plane j
    // once that is written back.
    S reg[ITER];
    fetch(reg, js);
    put(reg, 0);
    fetch(reg, js + 1);
    put(reg, 1);
    T nu_c, nu_n = T(0);
    uint64_t mw_c, mw_n = 0;
    cell_pre(js, nu_c, mw_c);
    __syncthreads();
    int A = 0;
    for (int j = js; j <= je; j++) {
        const int B = A ^ 1;
        const bool next = j + 2 <= je + 1;
        if (next) fetch(reg, j + 2);
        if (j + 1 <= je) cell_pre(j + 1, nu_n, mw_n);
        collide(j, A, B, nu_c, mw_c);
        __syncthreads();
        store(A, j, je);
        if (next) put(reg, A);
        __syncthreads();
        A = B;
        nu_c = nu_n;
        mw_c = mw_n;
    }
    // plane je + 1: the words of the chunk's last plane (twist in y)
    store(A, je + 1, je);
}

// ------------------------------------------------------------------ Simple
// In-place collide / moments, or pull from src (the two-lattice form of the
// reference's in-place push: slot d of an interior cell c takes src[c-e_d][d]
// when c-e_d is interior and keeps src[c][d] otherwise).
template <typename PR, int L, int MODE>
__device__ __forceinline__ void simple_cell(const typename PR::S* src,
                                            typename PR::S* dst,   // equal to src for in-place modes
                                            typename PR::T* mout,  // moments output of the MOM modes
                                            const typename PR::T* __restrict__ F, typename PR::N* __restrict__ Nu,
                                            const uint64_t* __restrict__ fmask, const Grid& g,
                                            const CollideParams<typename PR::T>& P, float rho0, int i, int j,
                                            int k) {
    using T = typename PR::T;
    using D = typename PR::D;
    const uint32_t n = cell_index(g, i, j, k);
    const uint32_t qn = qcell(g, i, j, k);

    T v[kQ];
    if (MODE == MODE_FUSED || MODE == MODE_MOM_PULL) {
        const bool im = i >= 2, ip = i <= g.xg1 - 1;  // cell i-1 / i+1 interior
        const bool jm = j >= 2, jp = j <= g.yg1 - 1;
        const bool km = k >= 2, kp = k <= g.zg1 - 1;
#pragma unroll
        for (int d = 0; d < kQ; d++) {
            const int ex = dir_x(d), ey = dir_y(d), ez = dir_z(d);
            const bool ok = (ex > 0 ? im : ex < 0 ? ip : true) && (ey > 0 ? jm : ey < 0 ? jp : true) &&
                            (ez > 0 ? km : ez < 0 ? kp : true);
            const int32_t off = -(ex * (int32_t)g.qyz + ey * (int32_t)g.qz + ez);
            const uint64_t a = qaddr<L>(g, ok ? qn + (uint32_t)off : qn, d);
            if constexpr (PR::kShift && MODE == MODE_FUSED) v[d] = load_dev<PR>(src, a, d);
            else v[d] = load_pop<PR>(src, a, d, rho0);
        }
    } else {
#pragma unroll
        for (int d = 0; d < kQ; d++) {
            if constexpr (PR::kShift && MODE == MODE_COLL) v[d] = mout[qaddr<L>(g, qn, d)];
            else v[d] = load_pop<PR>(src, qaddr<L>(g, qn, d), d, rho0);
        }
    }

    if (MODE == MODE_MOM || MODE == MODE_MOM_PULL) {
        T m[kQ];
        es_moments<T, D>(v, m);
#pragma unroll
        for (int d = 0; d < kQ; d++) mout[qaddr<L>(g, qn, d)] = m[d];
        return;
    }

    T fx = T(0), fy = T(0), fz = T(0);
    if (mask_test(fmask, g, i, j, k)) {
        fx = F[3 * (uint64_t)n + 0];
        fy = F[3 * (uint64_t)n + 1];
        fz = F[3 * (uint64_t)n + 2];
    }
    T nuc = P.use_les ? load_nu<PR>(Nu, n) : T(0);
    T p[kQ];
    if constexpr (MODE == MODE_FUSED && PR::kShift) {
        es_collide_dev<HwOps>(v, rest_density(rho0), fx, fy, fz, nuc, P, p);
    } else if (MODE == MODE_FUSED) {
        T m[kQ];
        es_moments<T, D>(v, m);
        es_collide<T, D>(m, fx, fy, fz, nuc, P, p);
    } else {
        es_collide<T, D>(v, fx, fy, fz, nuc, P, p);
    }
    if (P.use_les) store_nu<PR>(Nu, n, nuc);
#pragma unroll
    for (int d = 0; d < kQ; d++) {
        if constexpr (PR::kShift && MODE == MODE_FUSED) store_dev<PR>(dst, qaddr<L>(g, qn, d), d, p[d]);
        else store_pop<PR>(dst, qaddr<L>(g, qn, d), d, p[d], rho0);
    }
}

template <typename PR, int L, int MODE>
__global__ __launch_bounds__(kRowBlock) void k_simple(const typename PR::S* src, typename PR::S* dst,
                                                   typename PR::T* mout, const typename PR::T* __restrict__ F,
                                                   typename PR::N* __restrict__ Nu,
                                                   const uint64_t* __restrict__ fmask, Grid g,
                                                   CollideParams<typename PR::T> P, uint32_t nkb, float rho0) {
    const RowBlock rb = row_block(g, layout_block<L>(), nkb, blockDim.x);
    const int k = rb.k0 + (int)threadIdx.x;
    if (k > g.zg1) return;
    simple_cell<PR, L, MODE>(src, dst, mout, F, Nu, fmask, g, P, rho0, rb.i, rb.j, k);
}

// ------------------------------------------------------------------ Simple pull, IJKL through LDS
// index of direction d among the five with e_y = +1
TDLB_HD constexpr int ypos_index(int d) {
    int n = 0;
    for (int e = 0; e < d; e++)
        if (dir_y(e) > 0) n++;
    return n;
}

// The two-lattice pull in IJKL (src -> dst), staged like k_esotwist_ijkl_lds:
// a workgroup owns a tile of RI x KT cells in (i, k) and marches along j with
// two slabs of src records in LDS, rows i0 - 1 .. i0 + RI and cells
// k0 - 1 .. k0 + KT (the +-1 halo; the ghost layer is real memory and the pull
// never uses it) of planes j and j + 1.  The five directions with e_y = +1 pull
// from plane j - 1: a lane reads them for its next cell while plane j - 1 is
// still in LDS and carries them in registers.  Each cell writes its whole
// record of dst, so no word of dst has two writers and the slabs are read-only.
template <typename PR, int RI, int KT>
__global__ __launch_bounds__(RI * KT) void k_simple_ijkl_lds(const typename PR::S* __restrict__ src,
                                                          typename PR::S* __restrict__ dst,
                                                          const typename PR::T* __restrict__ F,
                                                          typename PR::N* __restrict__ Nu,
                                                          const uint64_t* __restrict__ fmask, Grid g,
                                                          CollideParams<typename PR::T> P, int nti, int ntk,
                                                          int jchunk, float rho0) {
    using T = typename PR::T;
    using D = typename PR::D;
    using S = typename PR::S;
    static_assert(!PR::kShift, "whole populations only");
    constexpr int NR = RI + 2, NK = KT + 2, REC = NR * NK, NT = RI * KT;
    __shared__ S slab[2][REC * kSlabPad];

    const uint32_t b = blockIdx.x;
    const int tk = (int)(b % (uint32_t)ntk);
    const int ti = (int)((b / (uint32_t)ntk) % (uint32_t)nti);
    const int tj = (int)(b / ((uint32_t)ntk * (uint32_t)nti));
    const int i0 = 1 + ti * RI, k0 = 1 + tk * KT;
    const int ri = min(RI, g.xg1 - i0 + 1), rk = min(KT, g.zg1 - k0 + 1);
    const int js = 1 + tj * jchunk;
    if (js > g.yg1) return;
    const int je = min(g.yg1, js + jchunk - 1);
    const int tid = (int)threadIdx.x;

    // slab row r holds records k0 - 1 .. k0 + rk of row i0 - 1 + r: RW words
    // from the word offset of record (i0 - 1 + r, j, k0 - 1)
    constexpr int RW = NK * kQ, RC = (RW + NT - 1) / NT, ITER = NR * RC;
    int loff[RC];
    bool wok[RC];
#pragma unroll
    for (int c = 0; c < RC; c++) {
        const int wr = c * NT + tid, kk = wr / kQ, sl = wr - kk * kQ;
        wok[c] = wr < RW && kk <= rk + 1;
        loff[c] = kk * kSlabPad + sl;
    }
    auto row_base = [&](int r, int j) __attribute__((always_inline)) -> int64_t {
        return (int64_t)cell_index(g, i0 - 1 + r, j, k0 - 1) * kQ;
    };
    auto fetch = [&](S (&reg)[ITER], int j) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (r > ri + 1) break;
            const S* row = src + row_base(r, j);
#pragma unroll
            for (int c = 0; c < RC; c++) {
                int t = c * NT + tid;
                asm volatile("" : "+v"(t));  // not hoisted across the plane loop
                if (wok[c]) reg[r * RC + c] = row[t];
            }
        }
    };
    auto put = [&](const S (&reg)[ITER], int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (r > ri + 1) break;
#pragma unroll
            for (int c = 0; c < RC; c++)
                if (wok[c]) slab[buf][r * NK * kSlabPad + loff[c]] = reg[r * RC + c];
        }
    };

    const int r = tid / KT, kk = tid - (tid / KT) * KT;
    const bool active = r < ri && kk < rk;
    const int ci = i0 + (active ? r : 0), ck = k0 + (active ? kk : 0);
    const bool im = ci >= 2, ip = ci <= g.xg1 - 1, km = ck >= 2, kp = ck <= g.zg1 - 1;
    // the source c - e_d of direction d is interior in x and z
    auto xz_ok = [&](int d) __attribute__((always_inline)) {
        const int ex = dir_x(d), ez = dir_z(d);
        return (ex > 0 ? im : ex < 0 ? ip : true) && (ez > 0 ? km : ez < 0 ? kp : true);
    };
    // slab position of slot d of the source (i - e_x, k - e_z), or of the cell itself
    auto src_at = [&](int d, bool shift) __attribute__((always_inline)) -> int {
        const int dr = shift ? -dir_x(d) : 0, dk = shift ? -dir_z(d) : 0;
        return ((r + 1 + dr) * NK + kk + 1 + dk) * kSlabPad + d;
    };
    // the e_y = +1 values of the next plane's cell (its sources lie in plane buf),
    // used only where the source is interior
    constexpr int NY = 5;
    auto carry_from = [&](int buf, T (&cy)[NY]) __attribute__((always_inline)) {
#pragma unroll
        for (int d = 0; d < kQ; d++)
            if (dir_y(d) > 0) cy[ypos_index(d)] = slab[buf][src_at(d, true)];
    };
    auto cell_pre = [&](int j, T& nu, uint64_t& mw) __attribute__((always_inline)) {
        nu = P.use_les ? load_nu<PR>(Nu, cell_index(g, ci, j, ck)) : T(0);
        mw = fmask[mask_word_index(g, ci, j, ck)];
    };

    S reg[ITER];
    fetch(reg, js - 1);
    put(reg, 0);
    fetch(reg, js);
    put(reg, 1);
    __syncthreads();
    T cy[NY];
    carry_from(0, cy);
    if (js + 1 <= g.yg1 + 1) fetch(reg, js + 1);
    __syncthreads();
    put(reg, 0);
    T nu_c, nu_n = T(0);
    uint64_t mw_c, mw_n = 0;
    cell_pre(js, nu_c, mw_c);
    __syncthreads();
    int A = 1;  // plane j; A ^ 1 holds plane j + 1
    for (int j = js; j <= je; j++) {
        const int B = A ^ 1;
        const bool next = j + 2 <= je + 1;
        if (next) fetch(reg, j + 2);
        if (j + 1 <= je) cell_pre(j + 1, nu_n, mw_n);
        if (active) {
            const bool jm = j >= 2, jp = j <= g.yg1 - 1;
            T v[kQ];
#pragma unroll
            for (int d = 0; d < kQ; d++) {
                // the pull of simple_cell: the source when it is interior in every axis, else the cell itself
                const int ey = dir_y(d);
                const bool ok = xz_ok(d) && (ey > 0 ? jm : ey < 0 ? jp : true);
                if (!ok) v[d] = slab[A][src_at(d, false)];
                else if (ey > 0) v[d] = cy[ypos_index(d)];
                else if (ey < 0) v[d] = slab[B][src_at(d, true)];
                else v[d] = slab[A][src_at(d, true)];
            }
            carry_from(A, cy);  // v holds this cell's values; cy now the next cell's
            const uint32_t cn = cell_index(g, ci, j, ck);
            T fx = T(0), fy = T(0), fz = T(0);
            if ((mw_c >> (ck & 63)) & 1ull) {
                fx = F[3 * (uint64_t)cn + 0];
                fy = F[3 * (uint64_t)cn + 1];
                fz = F[3 * (uint64_t)cn + 2];
            }
            T m[kQ], p[kQ];
            es_moments<T, D>(v, m);
            es_collide<T, D>(m, fx, fy, fz, nu_c, P, p);
            if (P.use_les) store_nu<PR>(Nu, cn, nu_c);
            S* out = dst + (uint64_t)cn * kQ;
#pragma unroll
            for (int d = 0; d < kQ; d++) out[d] = p[d];
        }
        __syncthreads();
        if (next) put(reg, A);  // plane j + 2 takes the buffer of plane j
        __syncthreads();
        A = B;
        nu_c = nu_n;
        mw_c = mw_n;
    }
}

// ------------------------------------------------------------------ data movement on every cell

// push: dst[c][d] = src[c-e_d][d] if c-e_d is interior, else src[c][d] (all cells)
template <typename W, int L>
__global__ void k_push_all(const W* __restrict__ src, W* __restrict__ dst, Grid g) {
    for (uint64_t n = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; n < g.size;
         n += (uint64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        decode_cell(g, n, i, j, k);
        const uint32_t qc = qcell(g, i, j, k);
#pragma unroll
        for (int d = 0; d < kQ; d++) {
            const int si = i - dir_x(d), sj = j - dir_y(d), sk = k - dir_z(d);
            const uint32_t s = is_interior(g, si, sj, sk) ? qcell(g, si, sj, sk) : qc;
            dst[qaddr<L>(g, qc, d)] = src[qaddr<L>(g, s, d)];
        }
    }
}

// pull: interior dst[c][d] = src[c-e_d][d]; ghosts copied
template <typename W, int L>
__global__ void k_pull_all(const W* __restrict__ src, W* __restrict__ dst, Grid g) {
    for (uint64_t n = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; n < g.size;
         n += (uint64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        decode_cell(g, n, i, j, k);
        const bool in = is_interior(g, i, j, k);
        const uint32_t qc = qcell(g, i, j, k);
#pragma unroll
        for (int d = 0; d < kQ; d++) {
            const uint32_t s = in ? qcell(g, i - dir_x(d), j - dir_y(d), k - dir_z(d)) : qc;
            dst[qaddr<L>(g, qc, d)] = src[qaddr<L>(g, s, d)];
        }
    }
}

// One face of the bounce-back, in place: ghost slot d (5 inward directions) of a
// face cell takes slot opp(d) of the neighbour g+e_d.  Faces run in the
// reference's order as separate launches because face rims read ghosts of
// faces handled later.
template <typename W, int L>
__global__ void k_bb_face(W* __restrict__ Q, Grid g, int face) {
    const int axis = face >> 1;        // 0 i, 1 j, 2 k
    const int inward = (face & 1) ? -1 : 1;
    const int pos = (face & 1) ? (axis == 0 ? g.xg1 + 1 : axis == 1 ? g.yg1 + 1 : g.zg1 + 1) : 0;
    const int na = axis == 0 ? g.yg1 : g.xg1;                 // first free axis extent
    const int nb = axis == 2 ? g.yg1 : g.zg1;                 // second free axis extent
    const uint64_t total = (uint64_t)na * nb;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int a = (int)(t / nb) + 1, b = (int)(t % nb) + 1;
        int i, j, k;
        if (axis == 0) { i = pos; j = a; k = b; }
        else if (axis == 1) { i = a; j = pos; k = b; }
        else { i = a; j = b; k = pos; }
        const uint32_t c = qcell(g, i, j, k);
#pragma unroll
        for (int d = 0; d < kQ; d++) {
            const int comp = axis == 0 ? dir_x(d) : axis == 1 ? dir_y(d) : dir_z(d);
            if (comp != inward) continue;
            const uint32_t s = qcell(g, i + dir_x(d), j + dir_y(d), k + dir_z(d));
            Q[qaddr<L>(g, c, d)] = Q[qaddr<L>(g, s, opposite(d))];
        }
    }
}

// The 12 edge lines of the bounce-back (Boundary.hpp:41-132): on an edge the
// one direction pointing inward on both boundary axes takes opp(d) of g+e_d.
template <typename W, int L>
__global__ void k_bb_edges(W* __restrict__ Q, Grid g) {
    const int nmax = max(g.xg1, max(g.yg1, g.zg1));
    const uint64_t total = 12ull * nmax;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int line = (int)(t / nmax), s = (int)(t % nmax) + 1;
        const int X0 = g.xg1 + 1, Y0 = g.yg1 + 1, Z0 = g.zg1 + 1;
        int i, j, k, d;
        switch (line) {
            case 0: i = 0; j = 0; k = s; d = Q07; break;
            case 1: i = X0; j = 0; k = s; d = Q14; break;
            case 2: i = 0; j = Y0; k = s; d = Q13; break;
            case 3: i = X0; j = Y0; k = s; d = Q08; break;
            case 4: i = 0; k = 0; j = s; d = Q09; break;
            case 5: i = 0; k = Z0; j = s; d = Q15; break;
            case 6: i = X0; k = Z0; j = s; d = Q10; break;
            case 7: i = X0; k = 0; j = s; d = Q16; break;
            case 8: j = 0; k = 0; i = s; d = Q11; break;
            case 9: j = 0; k = Z0; i = s; d = Q17; break;
            case 10: j = Y0; k = Z0; i = s; d = Q12; break;
            default: j = Y0; k = 0; i = s; d = Q18; break;
        }
        const int lim = line < 4 ? g.zg1 : line < 8 ? g.yg1 : g.xg1;
        if (s > lim) continue;
        const uint32_t c = qcell(g, i, j, k);
        const uint32_t src = qcell(g, i + dir_x(d), j + dir_y(d), k + dir_z(d));
        Q[qaddr<L>(g, c, d)] = Q[qaddr<L>(g, src, opposite(d))];
    }
}

// Ghost cells enumerated face by face without repeats:
// i = 0 and i = xg0 planes, then j = 0 / yg0 with 1 <= i <= xg1, then k = 0 / zg0
// with 1 <= i <= xg1, 1 <= j <= yg1.
__device__ __forceinline__ bool ghost_cell(const Grid& g, uint64_t t, int& i, int& j, int& k) {
    const uint64_t A = (uint64_t)g.yz;                       // one i-plane
    const uint64_t B = (uint64_t)g.xg1 * g.zg;               // one j-plane without i ghosts
    const uint64_t C = (uint64_t)g.xg1 * g.yg1;              // one k-plane without i, j ghosts
    if (t < 2 * A) {
        i = t < A ? 0 : g.xg1 + 1;
        const uint64_t r = t % A;
        j = (int)(r / g.zg);
        k = (int)(r % g.zg);
        return true;
    }
    t -= 2 * A;
    if (t < 2 * B) {
        j = t < B ? 0 : g.yg1 + 1;
        const uint64_t r = t % B;
        i = (int)(r / g.zg) + 1;
        k = (int)(r % g.zg);
        return true;
    }
    t -= 2 * B;
    if (t < 2 * C) {
        k = t < C ? 0 : g.zg1 + 1;
        const uint64_t r = t % C;
        i = (int)(r / g.yg1) + 1;
        j = (int)(r % g.yg1) + 1;
        return true;
    }
    return false;
}
__host__ __device__ inline uint64_t ghost_count(const Grid& g) {
    return 2ull * g.yz + 2ull * (uint64_t)g.xg1 * g.zg + 2ull * (uint64_t)g.xg1 * g.yg1;
}

// Ghost slots the reference rewrites every step from interior values: the push
// into the ghost layer (c-e_d interior) and the bounce-back slots whose source
// g+e_d is interior.  Rim bounce-back slots (ghost source) and never-written
// slots are left as they are in dst.
template <typename PR, int L>
__global__ void k_ghost_fill(const typename PR::S* __restrict__ src, typename PR::T* __restrict__ dst, Grid g,
                             float rho0) {
    const uint64_t total = ghost_count(g);
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        if (!ghost_cell(g, t, i, j, k)) continue;
        const int bi = (i == 0) ? 1 : (i == g.xg1 + 1) ? -1 : 0;  // inward sign per boundary axis
        const int bj = (j == 0) ? 1 : (j == g.yg1 + 1) ? -1 : 0;
        const int bk = (k == 0) ? 1 : (k == g.zg1 + 1) ? -1 : 0;
        const int nb = (bi != 0) + (bj != 0) + (bk != 0);
        const uint32_t c = qcell(g, i, j, k);
#pragma unroll
        for (int d = 0; d < kQ; d++) {
            const int ex = dir_x(d), ey = dir_y(d), ez = dir_z(d);
            const int si = i - ex, sj = j - ey, sk = k - ez;
            if (is_interior(g, si, sj, sk)) {
                dst[qaddr<L>(g, c, d)] = load_pop<PR>(src, qaddr<L>(g, qcell(g, si, sj, sk), d), d, rho0);
                continue;
            }
            bool bb = false;
            if (nb == 1) {
                bb = (bi != 0 && ex == bi) || (bj != 0 && ey == bj) || (bk != 0 && ez == bk);
            } else if (nb == 2) {
                bb = (bi == 0 || ex == bi) && (bj == 0 || ey == bj) && (bk == 0 || ez == bk) &&
                     (bi != 0 || ex == 0) && (bj != 0 || ey == 0) && (bk != 0 || ez == 0);
            }
            if (!bb) continue;
            const int ti = i + ex, tj = j + ey, tk = k + ez;
            if (!is_interior(g, ti, tj, tk)) continue;
            // opp(d) has the rest weight of d
            dst[qaddr<L>(g, c, d)] = load_pop<PR>(src, qaddr<L>(g, qcell(g, ti, tj, tk), opposite(d)), d, rho0);
        }
    }
}

template <typename W, int L>
__global__ void k_ghost_copy(const W* __restrict__ src, W* __restrict__ dst, Grid g) {
    const uint64_t total = ghost_count(g);
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        if (!ghost_cell(g, t, i, j, k)) continue;
        const uint32_t c = qcell(g, i, j, k);
#pragma unroll
        for (int d = 0; d < kQ; d++) dst[qaddr<L>(g, c, d)] = src[qaddr<L>(g, c, d)];
    }
}

// ------------------------------------------------------------------ F occupancy
// Bit per cell: any component of F[c] with a non-zero bit pattern.  gmask gets
// the ghost cells only.
template <typename T>
__global__ void k_mask_from_f(const T* __restrict__ F, uint64_t* __restrict__ mask, uint64_t* __restrict__ gmask,
                              Grid g) {
    const uint64_t words = (uint64_t)g.xg * g.yg * g.mask_w;
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < words;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t row = w / g.mask_w;
        const int kb = (int)(w % g.mask_w) * 64;
        const int i = (int)(row / g.yg), j = (int)(row % g.yg);
        uint64_t bits = 0, gbits = 0;
        for (int t = 0; t < 64; t++) {
            const int k = kb + t;
            if (k >= g.zg) break;
            const uint64_t n = cell_index(g, i, j, k);
            const T fx = F[3 * n], fy = F[3 * n + 1], fz = F[3 * n + 2];
            const bool nz = fx != T(0) || fy != T(0) || fz != T(0) || signbit(fx) || signbit(fy) || signbit(fz);
            if (nz) {
                bits |= 1ull << t;
                if (!is_interior(g, i, j, k)) gbits |= 1ull << t;
            }
        }
        mask[w] = bits;
        gmask[w] = gbits;
    }
}

// F as the reference holds it: the stored value where the bit is set, else 0.
template <typename T>
__global__ void k_masked_f(const T* __restrict__ F, const uint64_t* __restrict__ mask, T* __restrict__ out, Grid g) {
    for (uint64_t n = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; n < g.size;
         n += (uint64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        decode_cell(g, n, i, j, k);
        const bool on = mask_test(mask, g, i, j, k);
        out[3 * n + 0] = on ? F[3 * n + 0] : T(0);
        out[3 * n + 1] = on ? F[3 * n + 1] : T(0);
        out[3 * n + 2] = on ? F[3 * n + 2] : T(0);
    }
}

__global__ void k_zero_interior_o(uint8_t* __restrict__ O, Grid g) {
    for (uint64_t n = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; n < g.size;
         n += (uint64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        decode_cell(g, n, i, j, k);
        if (is_interior(g, i, j, k)) O[n] = 0;
    }
}

// ------------------------------------------------------------------ IB forcing
__device__ __forceinline__ int wrap_ik(int a, int a1) {  // Forcing.hpp:85-88
    if (a == 0) return a1;
    if (a == a1 + 1) return 1;
    return a;
}

// Before a forcing pass: F of a stencil cell that was not occupied after the
// previous pass is zero in the reference (F=0 where O==0), so clear it here,
// and mark the cell in the new occupancy mask.
template <typename T>
__device__ __forceinline__ void prepare_point(const PointsDev& p, int64_t q, T* __restrict__ F,
                                              const uint64_t* __restrict__ prev, uint64_t* __restrict__ next,
                                              const Grid& g) {
    const int pi = p.i[q], pj = p.j[q], pk = p.k[q];
    for (int k1 = -1; k1 <= 1; k1++)
        for (int i1 = -1; i1 <= 1; i1++) {
            const int i2 = wrap_ik(pi + i1, g.xg1), k2 = wrap_ik(pk + k1, g.zg1);
            const uint64_t c = cell_index(g, i2, pj, k2);
            if (!mask_test(prev, g, i2, pj, k2)) {
                F[3 * c + 0] = T(0);
                F[3 * c + 1] = T(0);
                F[3 * c + 2] = T(0);
            }
            atomicOr((unsigned long long*)&next[mask_word_index(g, i2, pj, k2)], 1ull << (k2 & 63));
        }
}

template <typename T>
__global__ void k_force_prepare(PointsDev p, int64_t begin, int64_t count, T* __restrict__ F,
                                const uint64_t* __restrict__ prev, uint64_t* __restrict__ next, Grid g) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= count) return;
    prepare_point(p, begin + t, F, prev, next, g);
}

// density and momentum of cell (i,j,k) as the reference's moments() would leave them
template <typename PR, int L, int STREAM, int MOMSRC, int EVEN>
__device__ __forceinline__ void cell_rho_j(const typename PR::S* __restrict__ Q, const Grid& g, int i, int j, int k,
                                           typename PR::T& rho, typename PR::T& jx, typename PR::T& jy,
                                           typename PR::T& jz, float rho0) {
    using T = typename PR::T;
    const uint32_t n = qcell(g, i, j, k);
    if (STREAM == StreamEsotwist) {
        const int32_t di = (i == g.xg1) ? -(int32_t)((uint32_t)(g.xg1 - 1) * g.qyz) : (int32_t)g.qyz;
        const int32_t dj = (j == g.yg1) ? -(int32_t)((uint32_t)(g.yg1 - 1) * g.qz) : (int32_t)g.qz;
        const int32_t dk = (k == g.zg1) ? -(g.zg1 - 1) : 1;
        if (MOMSRC == MOMSRC_STORED) {
            T m[4];
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const uint32_t h = n + (uint32_t)((twist_dx(d) ? di : 0) + (twist_dy(d) ? dj : 0) + (twist_dz(d) ? dk : 0));
                m[d] = (T)Q[qaddr<L>(g, h, EVEN ? d : opposite(d))];
            }
            rho = m[0]; jx = m[1]; jy = m[2]; jz = m[3];
        } else {
            T f[kQ];
#pragma unroll
            for (int d = 0; d < kQ; d++) {
                const uint32_t h = n + (uint32_t)((twist_dx(d) ? di : 0) + (twist_dy(d) ? dj : 0) + (twist_dz(d) ? dk : 0));
                f[d] = load_pop<PR>(Q, qaddr<L>(g, h, EVEN ? d : opposite(d)), d, rho0);
            }
            es_rho_j(f, rho, jx, jy, jz);
        }
    } else {
        if (MOMSRC == MOMSRC_STORED) {
            rho = (T)Q[qaddr<L>(g, n, 0)];
            jx = (T)Q[qaddr<L>(g, n, 1)];
            jy = (T)Q[qaddr<L>(g, n, 2)];
            jz = (T)Q[qaddr<L>(g, n, 3)];
        } else {
            const bool im = i >= 2, ip = i <= g.xg1 - 1;
            const bool jm = j >= 2, jp = j <= g.yg1 - 1;
            const bool km = k >= 2, kp = k <= g.zg1 - 1;
            T f[kQ];
#pragma unroll
            for (int d = 0; d < kQ; d++) {
                const int ex = dir_x(d), ey = dir_y(d), ez = dir_z(d);
                const bool ok = (ex > 0 ? im : ex < 0 ? ip : true) && (ey > 0 ? jm : ey < 0 ? jp : true) &&
                                (ez > 0 ? km : ez < 0 ? kp : true);
                const int32_t off = -(ex * (int32_t)g.qyz + ey * (int32_t)g.qz + ez);
                f[d] = load_pop<PR>(Q, qaddr<L>(g, ok ? n + (uint32_t)off : n, d), d, rho0);
            }
            es_rho_j(f, rho, jx, jy, jz);
        }
    }
}

// One level of the forcing: points of a level have pairwise disjoint stencils,
// so each thread runs the reference's gather / correct / spread for one point
// (Forcing.hpp:73-135) with the F values left by all earlier levels.
template <typename PR, int L, int STREAM, int MOMSRC, int EVEN>
__global__ __launch_bounds__(kBlock) void k_force_level(PointsDev p, int64_t begin, int64_t count,
                                                        typename PR::T* __restrict__ F,
                                                        const typename PR::S* __restrict__ Q, Grid g,
                                                        typename PR::T alfa, typename PR::T beta, float rho0) {
    using T = typename PR::T;
    using D = typename PR::D;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int64_t q = begin + t;
    const int pi = p.i[q], pj = p.j[q], pk = p.k[q];
    const T ifr = static_cast<const T*>(p.ifrac)[q], kfr = static_cast<const T*>(p.kfrac)[q];
    T wx[3], wz[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        wx[a] = ib_weight<T>(-ifr + T(a - 1));
        wz[a] = ib_weight<T>(-kfr + T(a - 1));
    }
    T rhoSum = T(0.0), xSum = T(0.0), ySum = T(0.0), zSum = T(0.0);
#pragma unroll
    for (int k1 = -1; k1 <= 1; k1++)
#pragma unroll
        for (int i1 = -1; i1 <= 1; i1++) {
            const int i2 = wrap_ik(pi + i1, g.xg1), k2 = wrap_ik(pk + k1, g.zg1);
            T rho, jx, jy, jz;
            cell_rho_j<PR, L, STREAM, MOMSRC, EVEN>(Q, g, i2, pj, k2, rho, jx, jy, jz, rho0);
            const uint64_t c = cell_index(g, i2, pj, k2);
            const T ppp = wx[i1 + 1] * wz[k1 + 1];
            const T xx = (T)((D)jx + D(0.5) * (D)F[3 * c + 0]);
            const T yy = (T)((D)jy + D(0.5) * (D)F[3 * c + 1]);
            const T zz = (T)((D)jz + D(0.5) * (D)F[3 * c + 2]);
            rhoSum += ppp * rho;
            xSum += ppp * xx;
            ySum += ppp * yy;
            zSum += ppp * zz;
        }
    xSum -= rhoSum * static_cast<const T*>(p.u)[q];
    ySum -= rhoSum * static_cast<const T*>(p.v)[q];
    zSum -= rhoSum * static_cast<const T*>(p.w)[q];
#pragma unroll
    for (int k1 = -1; k1 <= 1; k1++)
#pragma unroll
        for (int i1 = -1; i1 <= 1; i1++) {
            const int i2 = wrap_ik(pi + i1, g.xg1), k2 = wrap_ik(pk + k1, g.zg1);
            const uint64_t c = cell_index(g, i2, pj, k2);
            const T bp = beta * (wx[i1 + 1] * wz[k1 + 1]);
            F[3 * c + 0] = alfa * F[3 * c + 0] - bp * xSum;
            F[3 * c + 1] = alfa * F[3 * c + 1] - bp * ySum;
            F[3 * c + 2] = alfa * F[3 * c + 2] - bp * zSum;
        }
}

// ------------------------------------------------------------------ IB forcing with a moments cache
// Density and momentum of the stencil cells are computed once per forcing pass
// (they do not change while F is updated), so the level kernels issue only
// independent cache / F loads instead of 18 scattered population loads per
// stencil cell.
template <typename T>
struct alignas(4 * sizeof(T)) Mom4 {
    T r, x, y, z;
};

template <typename PR, int L, int STREAM, int MOMSRC, int EVEN>
__global__ __launch_bounds__(kBlock) void k_ib_moments_cells(const uint32_t* __restrict__ cells, int64_t n,
                                                             Mom4<typename PR::T>* __restrict__ cache,
                                                             const typename PR::S* __restrict__ Q, Grid g,
                                                             float rho0) {
    using T = typename PR::T;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t c = cells[t];
    const int i = (int)(c / g.yz);
    const uint32_t r = c - (uint32_t)i * g.yz;
    const int j = (int)(r / (uint32_t)g.zg);
    const int k = (int)(r - (uint32_t)j * (uint32_t)g.zg);
    Mom4<T> m;
    cell_rho_j<PR, L, STREAM, MOMSRC, EVEN>(Q, g, i, j, k, m.r, m.x, m.y, m.z, rho0);
    cache[t] = m;
}

template <typename PR, int L, int STREAM, int MOMSRC, int EVEN>
__global__ __launch_bounds__(kBlock) void k_ib_moments_points(PointsDev p, int64_t begin, int64_t count,
                                                              Mom4<typename PR::T>* __restrict__ cache,
                                                              const typename PR::S* __restrict__ Q, Grid g,
                                                              float rho0) {
    using T = typename PR::T;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= count * 9) return;
    const int64_t q = begin + t / 9;
    const int s = (int)(t % 9);
    const int i2 = wrap_ik(p.i[q] + (s % 3) - 1, g.xg1), k2 = wrap_ik(p.k[q] + (s / 3) - 1, g.zg1);
    Mom4<T> m;
    cell_rho_j<PR, L, STREAM, MOMSRC, EVEN>(Q, g, i2, p.j[q], k2, m.r, m.x, m.y, m.z, rho0);
    cache[t] = m;
}

// Same arithmetic as k_force_level, one point per 3 lanes of a wavefront: lane
// c of the group holds the stencil row k1 = c - 1, its cells s = 3c + m for
// i1 = m - 1 (s = (k1+1)*3 + (i1+1)), so the 9 gathers of a point run in
// parallel, three per lane.  Stencil cell s reads cache[sidx[9q+s]] (INDEXED)
// or cache[9(q-cache_base)+s].  Every lane of the group adds the 9 terms in
// the reference's order (s = 0..8, each product rounded before the add), so
// the sums are those of the sequential loop bit for bit.  The 9 cells of a
// point are distinct (the cached path needs 3 interior cells in x and z) and
// the points of a level have disjoint stencils, so each F entry has one writer.
// Three lanes per point put 21 points in a wavefront (7 with a lane per cell):
// a level's passes, each a chain of dependent loads, are a third as many.
constexpr int kGroup = 3;                    // lanes per point
constexpr int kGroupsPerWave = 64 / kGroup;  // 21 points per wavefront, lane 63 idle

// The inputs of stencil row c of point q, as loaded (force_point9 in two
// parts, so that a pass can load the next pass's points while it computes).
template <typename T>
struct PointRaw {
    int pi, pj, pk;
    T ifr, kfr, u, v, w;
    int64_t e[3];  // cache entries of the row's cells
};
template <typename T, bool INDEXED>
__device__ __forceinline__ PointRaw<T> point_raw(const PointsDev& p, int64_t q, bool valid, int c,
                                                 const int32_t* __restrict__ sidx, int64_t cache_base) {
    PointRaw<T> r{};
    if (valid) {
        r.pi = p.i[q];
        r.pj = p.j[q];
        r.pk = p.k[q];
        r.ifr = static_cast<const T*>(p.ifrac)[q];
        r.kfr = static_cast<const T*>(p.kfrac)[q];
        r.u = static_cast<const T*>(p.u)[q];
        r.v = static_cast<const T*>(p.v)[q];
        r.w = static_cast<const T*>(p.w)[q];
#pragma unroll
        for (int m = 0; m < 3; m++)
            r.e[m] = INDEXED ? (int64_t)sidx[9 * q + 3 * c + m] : 9 * (q - cache_base) + 3 * c + m;
    }
    return r;
}

template <typename T, typename D>
__device__ __forceinline__ void force_point9_raw(const PointRaw<T>& r, bool valid, int c, int base,
                                                 const Mom4<T>* __restrict__ cache, T* __restrict__ F, const Grid& g,
                                                 T alfa, T beta) {
    T w[3], tr[3], tx[3], ty[3], tz[3], fo[3][3];
    uint64_t cell[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        w[m] = tr[m] = tx[m] = ty[m] = tz[m] = fo[m][0] = fo[m][1] = fo[m][2] = T(0);
        cell[m] = 0;
    }
    if (valid) {
        const T wk = ib_weight<T>(-r.kfr + T(c - 1));
        const int k2 = wrap_ik(r.pk + c - 1, g.zg1);
        Mom4<T> mo[3];
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const int i2 = wrap_ik(r.pi + m - 1, g.xg1);
            cell[m] = cell_index(g, i2, r.pj, k2);
            mo[m] = cache[r.e[m]];
            fo[m][0] = F[3 * cell[m] + 0];
            fo[m][1] = F[3 * cell[m] + 1];
            fo[m][2] = F[3 * cell[m] + 2];
        }
#pragma unroll
        for (int m = 0; m < 3; m++) {
            w[m] = ib_weight<T>(-r.ifr + T(m - 1)) * wk;
            const T xx = (T)((D)mo[m].x + D(0.5) * (D)fo[m][0]);
            const T yy = (T)((D)mo[m].y + D(0.5) * (D)fo[m][1]);
            const T zz = (T)((D)mo[m].z + D(0.5) * (D)fo[m][2]);
            tr[m] = w[m] * mo[m].r;
            tx[m] = w[m] * xx;
            ty[m] = w[m] * yy;
            tz[m] = w[m] * zz;
        }
    }
    // terms s = 3 c2 + m in the order s = 0..8: row c2 from lane base + c2
    T rhoSum = T(0.0), xSum = T(0.0), ySum = T(0.0), zSum = T(0.0);
#pragma unroll
    for (int c2 = 0; c2 < kGroup; c2++)
#pragma unroll
        for (int m = 0; m < 3; m++) {
            rhoSum += __shfl(tr[m], base + c2, 64);
            xSum += __shfl(tx[m], base + c2, 64);
            ySum += __shfl(ty[m], base + c2, 64);
            zSum += __shfl(tz[m], base + c2, 64);
        }
    if (!valid) return;
    xSum -= rhoSum * r.u;
    ySum -= rhoSum * r.v;
    zSum -= rhoSum * r.w;
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const T bp = beta * w[m];
        F[3 * cell[m] + 0] = alfa * fo[m][0] - bp * xSum;
        F[3 * cell[m] + 1] = alfa * fo[m][1] - bp * ySum;
        F[3 * cell[m] + 2] = alfa * fo[m][2] - bp * zSum;
    }
}

// force_point9_raw with the row's force values in LDS: entry e of the
// plane's static cells at lds[3 (e - cb)], the same arithmetic in the same order.
template <typename T, typename D>
__device__ __forceinline__ void force_point9_lds(const PointRaw<T>& r, bool valid, int c, int base,
                                                 const Mom4<T>* __restrict__ cache, T* lds, int64_t cb, T alfa,
                                                 T beta) {
    T w[3], tr[3], tx[3], ty[3], tz[3], fo[3][3];
    int li[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        w[m] = tr[m] = tx[m] = ty[m] = tz[m] = fo[m][0] = fo[m][1] = fo[m][2] = T(0);
        li[m] = 0;
    }
    if (valid) {
        const T wk = ib_weight<T>(-r.kfr + T(c - 1));
        Mom4<T> mo[3];
#pragma unroll
        for (int m = 0; m < 3; m++) {
            li[m] = 3 * (int)(r.e[m] - cb);
            mo[m] = cache[r.e[m]];
            fo[m][0] = lds[li[m] + 0];
            fo[m][1] = lds[li[m] + 1];
            fo[m][2] = lds[li[m] + 2];
        }
#pragma unroll
        for (int m = 0; m < 3; m++) {
            w[m] = ib_weight<T>(-r.ifr + T(m - 1)) * wk;
            const T xx = (T)((D)mo[m].x + D(0.5) * (D)fo[m][0]);
            const T yy = (T)((D)mo[m].y + D(0.5) * (D)fo[m][1]);
            const T zz = (T)((D)mo[m].z + D(0.5) * (D)fo[m][2]);
            tr[m] = w[m] * mo[m].r;
            tx[m] = w[m] * xx;
            ty[m] = w[m] * yy;
            tz[m] = w[m] * zz;
        }
    }
    T rhoSum = T(0.0), xSum = T(0.0), ySum = T(0.0), zSum = T(0.0);
#pragma unroll
    for (int c2 = 0; c2 < kGroup; c2++)
#pragma unroll
        for (int m = 0; m < 3; m++) {
            rhoSum += __shfl(tr[m], base + c2, 64);
            xSum += __shfl(tx[m], base + c2, 64);
            ySum += __shfl(ty[m], base + c2, 64);
            zSum += __shfl(tz[m], base + c2, 64);
        }
    if (!valid) return;
    xSum -= rhoSum * r.u;
    ySum -= rhoSum * r.v;
    zSum -= rhoSum * r.w;
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const T bp = beta * w[m];
        lds[li[m] + 0] = alfa * fo[m][0] - bp * xSum;
        lds[li[m] + 1] = alfa * fo[m][1] - bp * ySum;
        lds[li[m] + 2] = alfa * fo[m][2] - bp * zSum;
    }
}

template <typename T, typename D, bool INDEXED>
__device__ __forceinline__ void force_point9(const PointsDev& p, int64_t q, bool valid, int c, int base,
                                             const int32_t* __restrict__ sidx, int64_t cache_base,
                                             const Mom4<T>* __restrict__ cache, T* __restrict__ F, const Grid& g,
                                             T alfa, T beta) {
    const PointRaw<T> r = point_raw<T, INDEXED>(p, q, valid, c, sidx, cache_base);
    force_point9_raw<T, D>(r, valid, c, base, cache, F, g, alfa, beta);
}

// lane -> (point index within the launch, stencil row, group base lane)
__device__ __forceinline__ int64_t group_point(int& c, int& base) {
    const int lane = (int)(threadIdx.x & 63), grp = lane / kGroup;
    c = lane - grp * kGroup;
    base = grp * kGroup;
    const int64_t wave = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
    return grp < kGroupsPerWave ? wave * kGroupsPerWave + grp : -1;
}

inline uint32_t group_blocks(int64_t points, int bs) {
    const int64_t waves = (points + kGroupsPerWave - 1) / kGroupsPerWave;
    const int64_t per_block = bs / 64;
    return (uint32_t)((waves + per_block - 1) / per_block);
}

template <typename T, typename D, bool INDEXED>
__global__ __launch_bounds__(kBlock) void k_force_level_cached(PointsDev p, int64_t begin, int64_t count,
                                                               const int32_t* __restrict__ sidx, int64_t cache_base,
                                                               const Mom4<T>* __restrict__ cache, T* __restrict__ F,
                                                               Grid g, T alfa, T beta) {
    int s, base;
    const int64_t t = group_point(s, base);
    const bool valid = t >= 0 && t < count;
    force_point9<T, D, INDEXED>(p, begin + (valid ? t : 0), valid, s, base, sidx, cache_base, cache, F, g, alfa,
                                beta);
}

// ------------------------------------------------------------------ IB forcing by planes
// A bound tank's forcing pass as independent plane tasks (ForceGeom in
// kernels.h).  A task first clears F and sets the occupancy bit of the blade
// stencil cells missing from the previous mask (prepare_point) and computes the
// density and momentum of its static and blade stencil cells; then it runs its
// static levels and its blade colour classes in order, passes of up to
// kPlanePoints points per workgroup, 3 lanes per point (force_point9), with a
// barrier after every level: the sequential list order of Forcing.hpp:73-135
// within the plane, without a kernel launch per level.  A plane with many
// points runs on a team of workgroups that meet at a barrier in global memory
// (team_barrier); teams are the first workgroups of the launch, so they are
// dispatched, and resident, before any other workgroup of it.
constexpr int kPlaneBlock = kForcePlaneBlock;
constexpr int kPlanePoints = (kPlaneBlock / 64) * kGroupsPerWave;
static_assert(kPlanePoints == kForcePlanePoints, "points per pass of a plane task");
// polls of a team barrier before it gives up and flags a time-out (about a
// second; every member of a team is resident, so a healthy barrier takes
// microseconds)
constexpr uint32_t kBarSpinLimit = 1u << 20;

// Sense-reversing barrier of a team of workgroups: bar[0] counts arrivals,
// bar[1] is the generation, bar[2] the time-out flag.  The last arrival resets
// the count (waiting for that reset to complete) and then advances the
// generation, so the barrier is ready for the next level and the next launch.
// Every wave waits for its own stores before the workgroup barrier that
// precedes the arrival: the barrier itself does not wait for them, and the
// arriving thread's fence covers only its own wave's stores, so a store of
// another wave could otherwise still be on its way to another L2 channel when
// the arrival is seen.  Local teams (every member on one XCD,
// ForceGeom::local_teams) exchange F and the caches through that XCD's L2, and
// each CU drops its L1 after the wait, with an invalidate at device scope (at
// work-group scope it leaves the L1 alone when work-groups are not split across
// CUs, and a member would read back F lines it cached before another member
// wrote them); the last arrival checks that every member reported the same XCC
// id (bar[2] bit 2 otherwise).  Other teams fence at agent scope (L2
// write-back before the arrival, invalidation after the wait).  The wait polls
// with relaxed loads (an acquire load would invalidate the L2 on every poll).
__device__ __forceinline__ void team_barrier(uint32_t* bar, int team, bool local) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        if (!local) __threadfence();
        if (local) {
            // bar[3] collects the XCC ids of the arrivals: one XCD must run the team
            uint32_t x;
            asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(x));
            (void)__hip_atomic_fetch_or(&bar[3], 1u << (x & 15u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
        const uint32_t gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (uint32_t)team - 1) {
            if (local) {
                const uint32_t seen = __hip_atomic_exchange(&bar[3], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (__builtin_popcount(seen) != 1) atomicOr(&bar[2], 2u);
            }
            __hip_atomic_exchange(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (++spins >= kBarSpinLimit) {
                    atomicOr(&bar[2], 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        if (!local) __threadfence();
    }
    __syncthreads();
    if (local) asm volatile("buffer_inv sc1" ::: "memory");
}

// lane -> point slot of a pass (wave-major, 21 groups of 3 lanes per wave), -1 for lane 63
__device__ __forceinline__ int plane_slot(int& c, int& base) {
    const int lane = (int)(threadIdx.x & 63), grp = lane / kGroup;
    c = lane - grp * kGroup;
    base = grp * kGroup;
    return grp < kGroupsPerWave ? (int)(threadIdx.x >> 6) * kGroupsPerWave + grp : -1;
}

// Blade colour class c of the generator's colour-major list (rushton.h
// blade_decode): its first list index, blades and radial points; point
// (blade bi, plane bp, radial ri) of the class is at base + (bi * planes + bp) * nr + ri.
struct BladeClass {
    int64_t base;
    int nb, nr;
};
__device__ __forceinline__ BladeEnum blade_enum_of(const BladeLayout& bl) {
    BladeEnum e;
    e.n_blades = bl.n_blades;
    e.n_planes = bl.n_planes;
    e.n_radial = bl.n_radial;
    e.n_bcol = bl.n_bcol;
    return e;
}
__device__ __forceinline__ BladeClass blade_class(const BladeEnum& e, int c, int64_t base) {
    const int bcol = c / 5, rmod = c % 5;
    BladeClass bc;
    bc.base = base;
    bc.nb = e.n_blades > bcol ? (e.n_blades - bcol + e.n_bcol - 1) / e.n_bcol : 0;
    bc.nr = e.n_radial > rmod ? (e.n_radial - rmod + 4) / 5 : 0;
    return bc;
}
// blade plane index of plane j (ghost included) and its colour classes (0: no blades)
__device__ __forceinline__ int blade_plane(const ForceGeom& fg, int j, int& ncls) {
    const BladeLayout& bl = fg.blade;
    const int bp = j - bl.ghost - bl.j_lo;
    ncls = (bl.n_blades > 0 && fg.n_rot > 0 && bp >= 0 && bp < bl.n_planes) ? 5 * blade_enum_of(bl).n_bcol : 0;
    return bp;
}

// Static stencil cells [c0, c1) of fg.cells: density and momentum into the cache,
// thread t of `threads` taking c0 + t, c0 + t + threads, ...
template <typename PR, int L, int STREAM, int MOMSRC, int EVEN>
__device__ __forceinline__ void gather_cells(const ForceGeom& fg, int64_t c0, int64_t c1, int64_t t,
                                             int64_t threads, const typename PR::S* __restrict__ Q, const Grid& g,
                                             float rho0) {
    auto* scache = static_cast<Mom4<typename PR::T>*>(fg.scache);
    for (int64_t e = c0 + t; e < c1; e += threads) {
        int i, j, k;
        decode_cell(g, fg.cells[e], i, j, k);
        Mom4<typename PR::T> m;
        cell_rho_j<PR, L, STREAM, MOMSRC, EVEN>(Q, g, i, j, k, m.r, m.x, m.y, m.z, rho0);
        scache[e] = m;
    }
}

// Blade stencil cells of plane j: occupancy bit in next, F = 0 where the
// previous mask lacks the cell (prepare_point), density and momentum into the
// blade cache.
template <typename PR, int L, int STREAM, int MOMSRC, int EVEN>
__device__ void gather_blades(const ForceGeom& fg, int j, int64_t t0, int64_t threads,
                              const typename PR::S* __restrict__ Q, typename PR::T* __restrict__ F,
                              const uint64_t* __restrict__ prev, uint64_t* __restrict__ next, const Grid& g,
                              float rho0) {
    using T = typename PR::T;
    auto* rcache = static_cast<Mom4<T>*>(fg.rcache);
    int ncls;
    const int bp = blade_plane(fg, j, ncls);
    const BladeEnum be = blade_enum_of(fg.blade);
    int64_t base = 0;
    for (int c = 0; c < ncls; c++) {
        const BladeClass bc = blade_class(be, c, base);
        const int64_t n9 = 9 * (int64_t)bc.nb * bc.nr;
        for (int64_t t = t0; t < n9; t += threads) {
            const int64_t pt = t / 9;
            const int s = (int)(t - 9 * pt);
            const int bi = (int)(pt / bc.nr), ri = (int)(pt - (int64_t)bi * bc.nr);
            const int64_t e = bc.base + ((int64_t)bi * fg.blade.n_planes + bp) * bc.nr + ri;
            const int64_t q = fg.n_static + e;
            const int pj = fg.pts.j[q];
            const int i2 = wrap_ik(fg.pts.i[q] + (s % 3) - 1, g.xg1), k2 = wrap_ik(fg.pts.k[q] + (s / 3) - 1, g.zg1);
            if (!mask_test(prev, g, i2, pj, k2)) {
                const uint64_t cc = cell_index(g, i2, pj, k2);
                F[3 * cc + 0] = T(0);
                F[3 * cc + 1] = T(0);
                F[3 * cc + 2] = T(0);
            }
            atomicOr((unsigned long long*)&next[mask_word_index(g, i2, pj, k2)], 1ull << (k2 & 63));
            Mom4<T> m;
            cell_rho_j<PR, L, STREAM, MOMSRC, EVEN>(Q, g, i2, pj, k2, m.r, m.x, m.y, m.z, rho0);
            rcache[9 * e + s] = m;
        }
        base += blade_class_size(be, c);
    }
}

// The plane's static levels, then its blade colour classes, by workgroup
// `rank` of a team of `team` (passes of (blockDim.x / 64) * 7 points), with
// sync() after every level.
template <typename T, typename D, typename Sync>
__device__ void plane_levels(const ForceGeom& fg, const PlaneDesc& pd, int rank, int team, T* __restrict__ F,
                             const Grid& g, T alfa, T beta, Sync&& sync, T* lds = nullptr) {
    if constexpr (sizeof(T) > 4) lds = nullptr;  // FP64: the registers of the LDS form spill
    const auto* scache = static_cast<const Mom4<T>*>(fg.scache);
    const auto* rcache = static_cast<const Mom4<T>*>(fg.rcache);
    int row, gbase;
    const int slot = plane_slot(row, gbase);
    const int64_t per = (int64_t)(blockDim.x >> 6) * kGroupsPerWave;
    const int64_t p0 = (int64_t)rank * per, pstride = (int64_t)team * per;
    // Each pass loads the next pass's points (positions, fractions, velocities,
    // cache entries) before it computes: the point lists do not change during
    // the pass.  Every static level ends with one sync(), also a level this
    // workgroup has no pass in.
    const int nlev = pd.n_levels;
    auto lev_end = [&](int lv) { return fg.lev_off[pd.lev + lv + 1]; };
    auto first = [&](int lv) { return fg.lev_off[pd.lev + lv] + p0; };
    auto raw_at = [&](int lv, int64_t qq) {
        const bool ok = lv < nlev && slot >= 0 && qq + slot < lev_end(lv);
        return point_raw<T, true>(fg.pts, ok ? qq + slot : 0, ok, row, fg.sidx, 0);
    };
    int l = 0;
    while (l < nlev && first(l) >= lev_end(l)) {
        sync();
        l++;
    }
    int64_t q = l < nlev ? first(l) : 0;
    PointRaw<T> cur = raw_at(l, q);
    while (l < nlev) {
        // the next pass of this workgroup: later in level l, else the first of the next level that has one
        int nl = l;
        int64_t nq = q + pstride;
        if (nq >= lev_end(l)) {
            nl = l + 1;
            while (nl < nlev && first(nl) >= lev_end(nl)) nl++;
            nq = nl < nlev ? first(nl) : 0;
        }
        const PointRaw<T> nxt = raw_at(nl, nq);
        if (lds)
            force_point9_lds<T, D>(cur, slot >= 0 && q + slot < lev_end(l), row, gbase, scache, lds, pd.cell_begin,
                                   alfa, beta);
        else
            force_point9_raw<T, D>(cur, slot >= 0 && q + slot < lev_end(l), row, gbase, scache, F, g, alfa, beta);
        for (int k = l; k < nl; k++) sync();  // levels l .. nl - 1 are done
        l = nl;
        q = nq;
        cur = nxt;
    }
    if (lds) {
        // the static levels are done (each ended with a barrier): their force
        // values back to F before the blade classes read it
        for (int64_t e = threadIdx.x; e < pd.cell_end - pd.cell_begin; e += blockDim.x) {
            const uint64_t cc = fg.cells[pd.cell_begin + e];  // cell_index of the entry
            F[3 * cc + 0] = lds[3 * e + 0];
            F[3 * cc + 1] = lds[3 * e + 1];
            F[3 * cc + 2] = lds[3 * e + 2];
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }
    int ncls;
    const int bp = blade_plane(fg, pd.j, ncls);
    const BladeEnum be = blade_enum_of(fg.blade);
    int64_t base = 0;
    for (int c = 0; c < ncls; c++) {
        const BladeClass bc = blade_class(be, c, base);
        const int64_t n = (int64_t)bc.nb * bc.nr;
        for (int64_t u0 = p0; u0 < n; u0 += pstride) {
            const int64_t t = u0 + slot;
            const bool valid = slot >= 0 && t < n;
            int64_t qb = fg.n_static;
            if (valid) {
                const int bi = (int)(t / bc.nr), ri = (int)(t - (int64_t)bi * bc.nr);
                qb += bc.base + ((int64_t)bi * fg.blade.n_planes + bp) * bc.nr + ri;
            }
            force_point9<T, D, false>(fg.pts, qb, valid, row, gbase, nullptr, fg.n_static, rcache, F, g, alfa, beta);
        }
        sync();
        base += blade_class_size(be, c);
    }
}

// workgroup `rank` of the plane's team
template <typename PR, int L, int STREAM, int MOMSRC, int EVEN>
__device__ void force_plane(const ForceGeom& fg, const PlaneDesc& pd, int rank,
                            const typename PR::S* __restrict__ Q, typename PR::T* __restrict__ F,
                            const uint64_t* __restrict__ prev, uint64_t* __restrict__ next, const Grid& g,
                            typename PR::T alfa, typename PR::T beta, float rho0,
                            typename PR::T* lds_buf = nullptr, int64_t lds_cells = 0, int64_t lds_min = 0) {
    const int team = pd.team;
    uint32_t* bar = team > 1 ? fg.bar + kBarWords * pd.bar : nullptr;
    const int64_t t0 = (int64_t)rank * blockDim.x + threadIdx.x, threads = (int64_t)team * blockDim.x;
    const bool local = fg.local_teams != 0;
    auto sync = [&]() {
        if (team > 1) team_barrier(bar, team, local);
        else __syncthreads();
    };
    gather_cells<PR, L, STREAM, MOMSRC, EVEN>(fg, pd.cell_begin, pd.cell_end, t0, threads, Q, g, rho0);
    gather_blades<PR, L, STREAM, MOMSRC, EVEN>(fg, pd.j, t0, threads, Q, F, prev, next, g, rho0);
    sync();
    // A solo plane whose static cells fit keeps their force values in LDS
    // through the static levels (every stencil cell of a plane task lies in
    // its own plane, which no other task touches): each level then waits on
    // LDS instead of on F in HBM.  Loaded after the blade cells' zeroing.
    using T = typename PR::T;
    const int64_t nc = pd.cell_end - pd.cell_begin;
    T* lds = (sizeof(T) <= 4 && lds_buf && team == 1 && nc <= lds_cells && nc >= lds_min) ? lds_buf : nullptr;
    if (lds) {
        for (int64_t e = threadIdx.x; e < nc; e += blockDim.x) {
            const uint64_t cc = fg.cells[pd.cell_begin + e];  // cell_index of the entry
            lds[3 * e + 0] = F[3 * cc + 0];
            lds[3 * e + 1] = F[3 * cc + 1];
            lds[3 * e + 2] = F[3 * cc + 2];
        }
        __syncthreads();
    }
    plane_levels<T, typename PR::D>(fg, pd, rank, team, F, g, alfa, beta, sync, lds);
}

template <typename PR, int L, int STREAM, int MOMSRC, int EVEN>
__global__ __launch_bounds__(kPlaneBlock) void k_force_planes(ForceGeom fg, const typename PR::S* __restrict__ Q,
                                                              typename PR::T* __restrict__ F,
                                                              const uint64_t* __restrict__ prev,
                                                              uint64_t* __restrict__ next, Grid g,
                                                              typename PR::T alfa, typename PR::T beta, float rho0) {
    __shared__ float4 lds_f[kPlaneLdsBytes / sizeof(float4)];
    const PlaneBlock pb = fg.blocks[blockIdx.x];
    if (pb.plane < 0) return;
    using T = typename PR::T;
    force_plane<PR, L, STREAM, MOMSRC, EVEN>(fg, fg.planes[pb.plane], pb.rank, Q, F, prev, next, g, alfa, beta,
                                             rho0, reinterpret_cast<T*>(lds_f),
                                             (int64_t)(kPlaneLdsBytes / (3 * sizeof(T))), kPlaneLdsMinCells);
}

// The levels of every plane after the gathers of the one-launch step: a
// workgroup (or team) per plane runs its blade stencil cells (occupancy, F,
// moments) and then its static levels and blade classes.
constexpr int kLevelBlock = 1024;
template <typename PR, int EVEN>
__global__ __launch_bounds__(kLevelBlock) void k_force_levels(ForceGeom fg, const typename PR::S* __restrict__ Q,
                                                              float* __restrict__ F, const uint64_t* __restrict__ prev,
                                                              uint64_t* __restrict__ next, Grid g, float alfa,
                                                              float beta, float rho0) {
    const PlaneBlock pb = fg.blocks[blockIdx.x];
    if (pb.plane < 0) return;
    const PlaneDesc& pd = fg.planes[pb.plane];
    const int team = pd.team;
    uint32_t* bar = team > 1 ? fg.bar + kBarWords * pd.bar : nullptr;
    const bool local = fg.local_teams != 0;
    auto sync = [&]() {
        if (team > 1) team_barrier(bar, team, local);
        else __syncthreads();
    };
    gather_blades<PR, LayoutLIJK, StreamEsotwist, MOMSRC_ESO_POPS, EVEN>(
        fg, pd.j, (int64_t)pb.rank * blockDim.x + threadIdx.x, (int64_t)team * blockDim.x, Q, F, prev, next, g, rho0);
    sync();
    plane_levels<float, typename PR::D>(fg, pd, pb.rank, team, F, g, alfa, beta, sync);
}

// ------------------------------------------------------------------ pipelined Esotwist step
// PipeArgs in kernels.h.  Work items of XCD x, in queue order (nJ planes
// [jlo, jhi], R = nJ * nkb row blocks per i step):
//   A  for i = 1 .. xg1: the R rows (i, j) of the range; after the rows of i
//      the gather of i - 2 (interior planes jlo < j < jhi; a gather of i reads
//      what the rows of i - 1 .. i + 1 wrote, i wrapping periodically), then
//      the gathers of xg1 - 1, xg1 and 1
//   B  the edge planes jlo and jhi (their gathers read rows of the
//      neighbouring XCDs), in chunks, once those XCDs have released their rows
//   C  the forcing levels of the range's planes (PlaneDesc), teams first
// A row item also copies its rows' static occupancy words into next.  Items
// wait only for earlier items, so every wait ends; spins are bounded anyway
// (kBarSpinLimit, flag in the XCD's counters).
//
// Counters (uint32, zero at launch): XCD x at ctr[64 x]: +0 next item, +1 rows
// done, +2 gathers done, +3 rows released to the other XCDs, +4 time-out flag;
// rows done per i step at ctr[64 kXcds + x (xg1 + 2) + i].  The XCD-local
// counters are relaxed agent-scope atomics: the rows a gather reads were
// written through this XCD's L2, and each row's stores are complete
// (s_waitcnt) before its count.  The release of a range for the neighbours is
// an agent-scope fence pair.
__device__ __forceinline__ uint32_t xcc_id() {
    uint32_t x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(x));
    return x;
}
__device__ __forceinline__ uint32_t ld_ctr(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// thread 0 polls until every counter reaches its target (or the spin limit)
__device__ __forceinline__ void pipe_wait(const uint32_t* a, uint32_t ta, const uint32_t* b, uint32_t tb,
                                          const uint32_t* c, uint32_t tc, uint32_t* flag) {
    if (threadIdx.x == 0) {
        uint32_t spins = 0;
        while (ld_ctr(a) < ta || (b && ld_ctr(b) < tb) || (c && ld_ctr(c) < tc)) {
            if (++spins >= kBarSpinLimit) {
                atomicOr(flag, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
}
// every store of the workgroup is complete in L2, then thread 0 counts it
__device__ __forceinline__ void pipe_done(uint32_t* ctr) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// rows of one item of the pipe (consecutive j of one i step)
constexpr uint32_t kPipeRows = 8;

// Occupancy of the pipe: 8 waves per SIMD for the 16-bit formats (61
// registers); 5 for FP32, the most without spills (92 registers; the separate
// FP32 kernel runs 6 at 74).
template <typename PR>
constexpr int pipe_waves() { return std::is_same<typename PR::S, float>::value ? 5 : 8; }

template <typename PR, int EVEN, bool HINT>
__global__ __launch_bounds__(kRowBlock) __attribute__((amdgpu_waves_per_eu(pipe_waves<PR>(), 8)))
void k_eso_pipe(PipeArgs a, Grid g, CollideParams<float> P, uint32_t nkb,
                                                      float rho0, float alfa, float beta) {
    using S = typename PR::S;
    constexpr int L = LayoutLIJK;
    constexpr int STREAM = StreamEsotwist;
    constexpr int MS = MOMSRC_ESO_POPS;
    // the gathers read the populations of the next step's collision
    constexpr int GE = EVEN ? 0 : 1;
    __shared__ uint32_t s_item;
    __shared__ int64_t s_pre[258];
    __shared__ int32_t s_c0[256];
    const uint32_t x = xcc_id() % kXcds;
    uint32_t* cx = a.ctr + 64 * x;
    uint64_t* tx = reinterpret_cast<uint64_t*>(cx + 8);  // phase times (wall clock)
    uint32_t* rows_done = a.ctr + 64 * kXcds + (size_t)x * (g.xg1 + 2);
    int jlo, jhi;
    pipe_range(g.yg1, (int)x, jlo, jhi);
    const int nj = jhi - jlo + 1, ni = g.xg1;
    const uint32_t R = (uint32_t)nj * nkb;                // row blocks per i step
    const uint32_t RG = (R + kPipeRows - 1) / kPipeRows;  // row items per i step
    const uint32_t nA = (uint32_t)ni * (RG + 1) + 3;
    const int rw = g.xg1 + 2;
    const int32_t* rc = a.rowcell;
    const int64_t cl0 = rc[(size_t)jlo * rw + 1], cl1 = rc[(size_t)jlo * rw + ni + 1];
    const int64_t ch0 = rc[(size_t)jhi * rw + 1], ch1 = rc[(size_t)jhi * rw + ni + 1];
    const int64_t chunk = 4 * (int64_t)blockDim.x;
    const uint32_t nBl = (uint32_t)((cl1 - cl0 + chunk - 1) / chunk), nBh = (uint32_t)((ch1 - ch0 + chunk - 1) / chunk);
    const uint32_t nB = nBl + nBh;
    const uint32_t gathers = (uint32_t)ni + nB;
    const uint32_t total = nA + nB;
    S* Q = static_cast<S*>(a.Q);
    float* F = static_cast<float*>(a.F);
    auto* Nu = static_cast<typename PR::N*>(a.Nu);
    const int tid = (int)threadIdx.x;

    // gather of row i over the interior planes of the range
    auto gather_row = [&](int i) {
        const int ninner = nj - 2;
        for (int p = tid; p < ninner; p += blockDim.x) {
            const size_t r = (size_t)(jlo + 1 + p) * rw + i;
            s_c0[p] = rc[r];
            s_pre[p + 1] = rc[r + 1] - rc[r];
        }
        __syncthreads();
        if (tid == 0) {
            s_pre[0] = 0;
            for (int p = 0; p < ninner; p++) s_pre[p + 1] += s_pre[p];
        }
        __syncthreads();
        const int64_t cells = s_pre[ninner];
        for (int64_t t = tid; t < cells; t += blockDim.x) {
            int lo = 0, hi = ninner - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_pre[mid] <= t) lo = mid;
                else hi = mid - 1;
            }
            const int64_t e = s_c0[lo] + (t - s_pre[lo]);
            gather_cells<PR, L, STREAM, MS, GE>(a.fg, e, e + 1, 0, 1, Q, g, rho0);
        }
    };
    auto wait_rows = [&](int i) {
        const int im = i == 1 ? ni : i - 1, ip = i == ni ? 1 : i + 1;
        pipe_wait(rows_done + im, R, rows_done + i, R, rows_done + ip, R, cx + 4);
    };
    auto gather_done = [&]() {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (tid == 0 &&
            __hip_atomic_fetch_add(cx + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == gathers)
            tx[2] = tx[3] = (uint64_t)wall_clock64();
    };

    // the next item is taken while the current one runs (each waits only for earlier items)
    uint32_t ticket = 0;
    if (tid == 0) ticket = __hip_atomic_fetch_add(cx, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        __syncthreads();
        if (tid == 0) {
            s_item = ticket;
            if (ticket == 0) tx[0] = (uint64_t)wall_clock64();
            if (ticket < total) ticket = __hip_atomic_fetch_add(cx, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const uint32_t u = s_item;
        if (u >= total) break;
        if (u < (uint32_t)ni * (RG + 1)) {
            const int i = (int)(u / (RG + 1)) + 1;
            const uint32_t r = u % (RG + 1);
            if (r < RG) {
                // A: row blocks [r G, r G + G) of i step i: row (i, jlo + b / nkb), k chunk b % nkb
                const uint32_t b0 = r * kPipeRows, b1 = min(b0 + kPipeRows, R);
                for (uint32_t b = b0; b < b1; b++) {
                    const int j = jlo + (int)(b / nkb);
                    const int k = (int)((b % nkb) * blockDim.x) + 1 + tid;
                    if (k <= g.zg1)
                        esotwist_cell<PR, L, MODE_FUSED, EVEN, HINT>(Q, nullptr, F, Nu, a.prev, g, P, rho0, i, j, k);
                    if (b % nkb == 0) {
                        // static occupancy words of the row, and of the ghost rows beside it
                        const int ia = i == 1 ? 0 : i, ib = i == ni ? ni + 1 : i;
                        const int ja = j == 1 ? 0 : j, jb = j == g.yg1 ? g.yg1 + 1 : j;
                        const int nw = (int)g.mask_w;
                        const int cnt = (ib - ia + 1) * (jb - ja + 1) * nw;
                        for (int t = tid; t < cnt; t += blockDim.x) {
                            const int w = t % nw, rr = t / nw;
                            const int jj = ja + rr % (jb - ja + 1), ii = ia + rr / (jb - ja + 1);
                            const uint64_t wi =
                                ((uint64_t)ii * (uint64_t)g.yg + (uint64_t)jj) * g.mask_w + (uint64_t)w;
                            a.next[wi] = a.static_base[wi];
                        }
                    }
                }
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                __syncthreads();
                if (tid == 0) {
                    __hip_atomic_fetch_add(rows_done + i, b1 - b0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const uint32_t done =
                        __hip_atomic_fetch_add(cx + 1, b1 - b0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (done + (b1 - b0) == (uint32_t)ni * R) {
                        // every row of the range is in this XCD's L2: write it back for the neighbours
                        tx[1] = (uint64_t)wall_clock64();
                        __threadfence();
                        __hip_atomic_store(cx + 3, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            } else if (i >= 4) {
                wait_rows(i - 2);
                gather_row(i - 2);
                gather_done();
            }
        } else if (u < nA) {
            const uint32_t tail = u - (uint32_t)ni * (RG + 1);
            const int i = tail == 0 ? ni - 1 : tail == 1 ? ni : 1;
            wait_rows(i);
            gather_row(i);
            gather_done();
        } else if (u < nA + nB) {
            // B: the edge planes, after the neighbouring ranges are released
            const uint32_t b = u - nA;
            const uint32_t xl = (x + kXcds - 1) % kXcds, xh = (x + 1) % kXcds;
            pipe_wait(a.ctr + 64 * xl + 3, 1, cx + 3, 1, a.ctr + 64 * xh + 3, 1, cx + 4);
            __threadfence();
            const int64_t c0 = b < nBl ? cl0 + (int64_t)b * chunk : ch0 + (int64_t)(b - nBl) * chunk;
            const int64_t c1 = b < nBl ? cl1 : ch1;
            gather_cells<PR, L, STREAM, MS, GE>(a.fg, c0, c0 + chunk < c1 ? c0 + chunk : c1, tid, blockDim.x, Q, g,
                                                 rho0);
            gather_done();
        }
    }
}

__global__ void k_xcc_ids(uint32_t* out) {
    if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

// ------------------------------------------------------------------ FP64 sums and macroscopic fields
// Both read density and momentum through cell_rho_j, so they work on stored
// moments and on the populations of the fused path alike, without changing
// the state.  Interior cells are enumerated t = ((i-1)*y + (j-1))*z + (k-1).
__device__ __forceinline__ void interior_cell(const Grid& g, uint64_t t, int& i, int& j, int& k) {
    k = (int)(t % (uint64_t)g.zg1) + 1;
    const uint64_t r = t / (uint64_t)g.zg1;
    j = (int)(r % (uint64_t)g.yg1) + 1;
    i = (int)(r / (uint64_t)g.yg1) + 1;
}

template <typename T>
__device__ __forceinline__ void cell_force(const T* __restrict__ F, const uint64_t* __restrict__ fmask,
                                           const Grid& g, int i, int j, int k, T& fx, T& fy, T& fz) {
    fx = fy = fz = T(0);
    if (mask_test(fmask, g, i, j, k)) {
        const uint64_t c = cell_index(g, i, j, k);
        fx = F[3 * c + 0];
        fy = F[3 * c + 1];
        fz = F[3 * c + 2];
    }
}

// Per block: {sum rho, sum jx, jy, jz, sum Fx, Fy, Fz, torque} in FP64.  The
// grid is fixed, so the partial sums and their combination are deterministic.
template <typename PR, int L, int STREAM, int MOMSRC, int EVEN>
__global__ __launch_bounds__(kBlock) void k_reduce_cells(const typename PR::S* __restrict__ Q,
                                                         const typename PR::T* __restrict__ F,
                                                         const uint64_t* __restrict__ fmask, Grid g, ReduceArgs ra,
                                                         float rho0, double* __restrict__ partial) {
    using T = typename PR::T;
    double acc[kReduceVals];
#pragma unroll
    for (int v = 0; v < kReduceVals; v++) acc[v] = 0.0;
    const uint64_t n = (uint64_t)g.xg1 * g.yg1 * g.zg1;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_cell(g, t, i, j, k);
        T rho, jx, jy, jz, fx, fy, fz;
        cell_rho_j<PR, L, STREAM, MOMSRC, EVEN>(Q, g, i, j, k, rho, jx, jy, jz, rho0);
        cell_force(F, fmask, g, i, j, k, fx, fy, fz);
        acc[0] += (double)rho;
        acc[1] += (double)jx;
        acc[2] += (double)jy;
        acc[3] += (double)jz;
        acc[4] += (double)fx;
        acc[5] += (double)fy;
        acc[6] += (double)fz;
        const double x = (double)(i - ra.ghost) - ra.cx, z = (double)(k - ra.ghost) - ra.cz;
        if (x * x + z * z <= ra.r2) acc[7] += x * (double)fz - z * (double)fx;
    }
    __shared__ double red[kReduceVals][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < kReduceVals; v++) {
        double a = acc[v];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
        if (lane == 0) red[v][wave] = a;
    }
    __syncthreads();
    if (threadIdx.x < kReduceVals) {
        double a = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) a += red[threadIdx.x][w];
        partial[(uint64_t)blockIdx.x * kReduceVals + threadIdx.x] = a;
    }
}

__global__ void k_reduce_final(const double* __restrict__ partial, int nblocks, double* __restrict__ out) {
    if (threadIdx.x >= kReduceVals) return;
    double a = 0.0;
    for (int b = 0; b < nblocks; b++) a += partial[(uint64_t)b * kReduceVals + threadIdx.x];
    out[threadIdx.x] = a;
}

// {rho, ux, uy, uz} per interior cell, u as QVec::velocity(F) forms it
// (QVec.hpp:125-131: (1.0 / rho) * (j + 0.5 * F), double literals).
template <typename PR, int L, int STREAM, int MOMSRC, int EVEN>
__global__ __launch_bounds__(kBlock) void k_macroscopic(const typename PR::S* __restrict__ Q,
                                                        const typename PR::T* __restrict__ F,
                                                        const uint64_t* __restrict__ fmask, Grid g, float rho0,
                                                        typename PR::T* __restrict__ out) {
    using T = typename PR::T;
    using D = typename PR::D;
    const uint64_t n = (uint64_t)g.xg1 * g.yg1 * g.zg1;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_cell(g, t, i, j, k);
        T rho, jx, jy, jz, fx, fy, fz;
        cell_rho_j<PR, L, STREAM, MOMSRC, EVEN>(Q, g, i, j, k, rho, jx, jy, jz, rho0);
        cell_force(F, fmask, g, i, j, k, fx, fy, fz);
        const D inv = D(1.0) / (D)rho;
        out[4 * t + 0] = rho;
        out[4 * t + 1] = (T)(inv * ((D)jx + D(0.5) * (D)fx));
        out[4 * t + 2] = (T)(inv * ((D)jy + D(0.5) * (D)fy));
        out[4 * t + 3] = (T)(inv * ((D)jz + D(0.5) * (D)fz));
    }
}

// ------------------------------------------------------------------ regrid and output planes
// doubleResolutionFullCU (ComputeUnit.hpp:229-316, 480-502): every cell of the
// doubled grid, ghosts included, takes old cell (i/2 + i%2, j/2 + j%2, k/2 + k%2).
__device__ __forceinline__ uint32_t regrid_source(const Grid& og, const Grid& ng, uint64_t n) {
    int i, j, k;
    decode_cell(ng, n, i, j, k);
    return cell_index(og, i / 2 + (i & 1), j / 2 + (j & 1), k / 2 + (k & 1));
}

template <typename W, int L>
__global__ void k_regrid_q(const W* __restrict__ old, W* __restrict__ nw, Grid og, Grid ng) {
    for (uint64_t n = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; n < ng.size;
         n += (uint64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        decode_cell(ng, n, i, j, k);
        const uint32_t to = qcell(ng, i, j, k);
        const uint32_t from = qcell(og, i / 2 + (i & 1), j / 2 + (j & 1), k / 2 + (k & 1));
#pragma unroll
        for (int l = 0; l < kQ; l++) nw[qaddr<L>(ng, to, l)] = old[qaddr<L>(og, from, l)];
    }
}

template <typename W>
__global__ void k_regrid_cells(const W* __restrict__ old, W* __restrict__ nw, Grid og, Grid ng) {
    for (uint64_t n = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; n < ng.size;
         n += (uint64_t)gridDim.x * blockDim.x)
        nw[n] = old[regrid_source(og, ng, n)];
}

// savePlaneXZ<float, 4> (ComputeUnitOutput.hpp:354-447) for every plane cell
// (i-major); flags bit 0: a QVec record (cell not excluded), bit 1: an F3 record
// (Force::isNotZero: all three components non-zero, QVec.hpp:54-56).
template <typename T, int L>
__global__ void k_plane_xz(const T* __restrict__ Q, const T* __restrict__ F, const uint64_t* __restrict__ fmask,
                           const uint8_t* __restrict__ exclude, Grid g, int cut, int64_t i0, int64_t k0,
                           PlaneRec4* __restrict__ q, PlaneRec3* __restrict__ f, uint8_t* __restrict__ flags) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)g.xg1 * g.zg1) return;
    const int i = (int)(t / g.zg1) + 1, k = (int)(t % g.zg1) + 1;
    const uint32_t n = cell_index(g, i, cut, k);
    uint8_t fl = 0;
    if (!(exclude && exclude[n])) {
        fl = 1;
        PlaneRec4 r;
        r.i = (uint16_t)(i0 + i - 1);
        r.j = (uint16_t)(cut - 1);
        r.k = (uint16_t)(k0 + k - 1);
        r.pad = 0;
#pragma unroll
        for (int l = 0; l < 4; l++) r.q[l] = (float)Q[qaddr<L>(g, qcell(g, i, cut, k), l)];
        q[t] = r;
        if (mask_test(fmask, g, i, cut, k)) {
            const T fx = F[3 * (uint64_t)n], fy = F[3 * (uint64_t)n + 1], fz = F[3 * (uint64_t)n + 2];
            if (fx != T(0) && fy != T(0) && fz != T(0)) {
                fl |= 2;
                PlaneRec3 s;
                s.i = (uint16_t)(i0 + i - 1);
                s.j = (uint16_t)cut;
                s.k = (uint16_t)(k0 + k - 1);
                s.pad = 0;
                s.q[0] = (float)fx;
                s.q[1] = (float)fy;
                s.q[2] = (float)fz;
                f[t] = s;
            }
        }
    }
    flags[t] = fl;
}

// calcVorticityXZ / XY (ComputeUnitOutput.hpp:113-160, 198-245) before the
// picture: log|curl u|^2 from central differences of QVec::velocity() (j / rho,
// no force term) read through the streaming accessor; excluded cells NaN.  The
// twisted read of a far-side ghost cell leaves the array in the reference, so
// such neighbours give NaN here.  axis 0: plane j = cut, (x, z); 1: k = cut, (x, y).
template <typename PR, int L, int STREAM, int MOMSRC, int EVEN>
__global__ void k_vorticity(const typename PR::S* __restrict__ Q, Grid g, int axis, int cut,
                            const uint8_t* __restrict__ exclude, typename PR::T* __restrict__ out, float rho0) {
    using T = typename PR::T;
    const int nb = axis == 0 ? g.zg1 : g.yg1;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)g.xg1 * nb) return;
    const int i = (int)(t / nb) + 1, c = (int)(t % nb) + 1;
    const int j = axis == 0 ? cut : c, k = axis == 0 ? c : cut;
    if (exclude && exclude[cell_index(g, i, j, k)]) {
        out[t] = (T)NAN;
        return;
    }
    auto vel = [&](int a, int b, int e, T& ux, T& uy, T& uz) {
        if (STREAM == StreamEsotwist && (a == g.xg1 + 1 || b == g.yg1 + 1 || e == g.zg1 + 1)) {
            ux = uy = uz = (T)NAN;
            return;
        }
        T rho, jx, jy, jz;
        cell_rho_j<PR, L, STREAM, MOMSRC, EVEN>(Q, g, a, b, e, rho, jx, jy, jz, rho0);
        ux = jx / rho;
        uy = jy / rho;
        uz = jz / rho;
    };
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
    out[t] = (T)log((T)(uyzuzy * uyzuzy + uzxuxz * uzxuxz + uxyuyx * uxyuyx));
}

// ------------------------------------------------------------------ geometry on the device
template <typename T>
__device__ __forceinline__ void blade_gen_point(const BladeLayout& bl, const BladeStep& st, const PointsDev& out,
                                                int64_t offset, int64_t t) {
    BladeEnum e;
    e.n_blades = bl.n_blades;
    e.n_planes = bl.n_planes;
    e.n_radial = bl.n_radial;
    e.n_bcol = bl.n_bcol;
    int b, pl, r, cls;
    blade_decode(e, t, b, pl, r, cls);
    tdlb_ib_point pt;
    blade_point(bl.cx, bl.cz, bl.r_inner, bl.j_lo + pl, r, st.cosb[b], st.sinb[b], st.omega, &pt);
    const int64_t q = offset + t;
    out.i[q] = (int32_t)(pt.i + bl.ghost);
    out.j[q] = (int32_t)(pt.j + bl.ghost);
    out.k[q] = (int32_t)(pt.k + bl.ghost);
    static_cast<T*>(out.ifrac)[q] = (T)pt.i_frac;
    static_cast<T*>(out.kfrac)[q] = (T)pt.k_frac;
    static_cast<T*>(out.u)[q] = (T)pt.u_delta;
    static_cast<T*>(out.v)[q] = (T)pt.v_delta;
    static_cast<T*>(out.w)[q] = (T)pt.w_delta;
}

template <typename T>
__global__ void k_gen_blades(BladeLayout bl, BladeStep st, PointsDev out, int64_t offset, int64_t count) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= count) return;
    blade_gen_point<T>(bl, st, out, offset, t);
}

// getExternalPoints as a mask: a cell is outside when its centre lies beyond the
// tank radius from the axis (exact: integer and half-integer coordinates).
__global__ void k_solid_mask(double cx, double cz, double radius, uint8_t* __restrict__ out, Grid g, int ghost) {
    for (uint64_t n = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; n < g.size;
         n += (uint64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        decode_cell(g, n, i, j, k);
        const double dx = (double)(i - ghost) - cx, dz = (double)(k - ghost) - cz;
        out[n] = (dx * dx + dz * dz > radius * radius) ? 1 : 0;
    }
}

template <typename T>
CollideParams<T> to_params(const CollideArgs& c) {
    return make_collide_params<T>(T(c.nu), T(c.cs0), T(c.g3), c.use_les);
}
template <typename T, typename D>
CollideParams<T> to_params(const CollideArgs& c) {
    return make_collide_params<T, D>(T(c.nu), T(c.cs0), T(c.g3), c.use_les);
}

uint32_t grid_blocks(uint64_t n) {
    uint64_t b = (n + kBlock - 1) / kBlock;
    if (b > 65536ull * 16) b = 65536ull * 16;
    return (uint32_t)(b ? b : 1);
}

// rows of interior cells: nkb blocks of bk threads along k, bk a multiple of
// the wavefront (64) covering the row in as few blocks of at most bmax lanes
inline void row_launch(const Grid& g, uint32_t& nkb, uint32_t& nblocks, uint32_t& bk, int bmax) {
    nkb = (uint32_t)((g.zg1 + bmax - 1) / bmax);
    const uint32_t per = ((uint32_t)g.zg1 + nkb - 1) / nkb;
    bk = (per + 63u) / 64u * 64u;
    nblocks = nkb * (uint32_t)g.yg1 * (uint32_t)g.xg1;
}

// the LDS-staged IJKL step: tile RI x KT cells, about kIjklBlocks workgroups.
// Measured at 512^3 (FP32): 4 x 64 and 16384 workgroups 5.48 ms, 8 x 32 5.51 ms
// (6144 workgroups 5.66 ms), 16 x 16 7.03 ms, 8 x 16 6.24 ms.
constexpr int kIjklTileI = 4, kIjklTileK = 64, kIjklBlocks = 16384;
// the Simple pull through LDS: tile RI x KT cells (halo on both sides)
constexpr int kIjklPullI = 8, kIjklPullK = 32;
// TDLB_IJKL_LDS=0 selects the one-row kernel for IJKL (for comparison)
inline bool ijkl_lds() {
    static const bool on = [] {
        const char* v = std::getenv("TDLB_IJKL_LDS");
        return !(v && v[0] == '0');
    }();
    return on;
}

template <typename PR, int L>
hipError_t esotwist_t(int mode, int even, void* Q, void* mout, const void* F, void* Nu, const uint64_t* fm,
                      const Grid& g, const CollideArgs& c, float rho0, hipStream_t s) {
    using S = typename PR::S;
    using T = typename PR::T;
    uint32_t nkb, nb, bk;
    row_launch(g, nkb, nb, bk, row_block_max<S>());
    S* q = static_cast<S*>(Q);
    T* mo = static_cast<T*>(mout);
    const T* f = static_cast<const T*>(F);
    auto* nu = static_cast<typename PR::N*>(Nu);
    const auto P = to_params<T, typename PR::D>(c);
#define TDLB_ESO(M, E, H) \
    hipLaunchKernelGGL((k_esotwist<PR, L, M, E, H>), dim3(nb), dim3(bk), 0, s, q, mo, f, nu, fm, g, P, nkb, rho0)
    if (mode == MODE_COLL) { if (even) TDLB_ESO(MODE_COLL, 1, false); else TDLB_ESO(MODE_COLL, 0, false); }
    else if (mode == MODE_MOM) { if (even) TDLB_ESO(MODE_MOM, 1, false); else TDLB_ESO(MODE_MOM, 0, false); }
    else if (mode == MODE_FUSED) {
        if constexpr (L == LayoutLIJK) {
            if (line_rows<S>(g)) {
                if (even) TDLB_ESO(MODE_FUSED, 1, true); else TDLB_ESO(MODE_FUSED, 0, true);
                return hipGetLastError();
            }
        }
        if constexpr (L == LayoutIJKL && !PR::kShift && std::is_same<typename PR::T, float>::value) {
            if (ijkl_lds()) {
                constexpr int RI = kIjklTileI, KT = kIjklTileK;
                const int nti = (g.xg1 + RI - 1) / RI, ntk = (g.zg1 + KT - 1) / KT;
                // j chunks: enough workgroups to fill the GPU several times over
                const int cols = nti * ntk;
                const int want = (kIjklBlocks + cols - 1) / cols;
                const int jchunk = std::max(4, (g.yg1 + want - 1) / want);
                const int nch = (g.yg1 + jchunk - 1) / jchunk;
                const dim3 grid((uint32_t)(cols * nch)), block((uint32_t)(RI * KT));
                if (even)
                    hipLaunchKernelGGL((k_esotwist_ijkl_lds<PR, 1, RI, KT>), grid, block, 0, s, q, f, nu, fm, g, P, nti,
                                       ntk, jchunk, rho0);
                else
                    hipLaunchKernelGGL((k_esotwist_ijkl_lds<PR, 0, RI, KT>), grid, block, 0, s, q, f, nu, fm, g, P, nti,
                                       ntk, jchunk, rho0);
                return hipGetLastError();
            }
        }
        if (even) TDLB_ESO(MODE_FUSED, 1, false); else TDLB_ESO(MODE_FUSED, 0, false);
    }
    else return hipErrorInvalidValue;
#undef TDLB_ESO
    return hipGetLastError();
}

template <typename PR, int L>
hipError_t simple_t(int mode, const void* src, void* dst, void* mout, const void* F, void* Nu, const uint64_t* fm,
                    const Grid& g, const CollideArgs& c, float rho0, hipStream_t s) {
    using S = typename PR::S;
    using T = typename PR::T;
    uint32_t nkb, nb, bk;
    row_launch(g, nkb, nb, bk, row_block_max<S>());
    const S* a = static_cast<const S*>(src);
    S* b = static_cast<S*>(dst);
    T* mo = static_cast<T*>(mout);
    const T* f = static_cast<const T*>(F);
    auto* nu = static_cast<typename PR::N*>(Nu);
    const auto P = to_params<T, typename PR::D>(c);
    if constexpr (L == LayoutIJKL && !PR::kShift && std::is_same<typename PR::T, float>::value) {
        if (mode == MODE_FUSED && ijkl_lds()) {
            constexpr int RI = kIjklPullI, KT = kIjklPullK;
            const int nti = (g.xg1 + RI - 1) / RI, ntk = (g.zg1 + KT - 1) / KT;
            const int cols = nti * ntk;
            const int want = (kIjklBlocks + cols - 1) / cols;
            const int jchunk = std::max(4, (g.yg1 + want - 1) / want);
            const int nch = (g.yg1 + jchunk - 1) / jchunk;
            hipLaunchKernelGGL((k_simple_ijkl_lds<PR, RI, KT>), dim3((uint32_t)(cols * nch)), dim3((uint32_t)(RI * KT)),
                               0, s, a, b, f, nu, fm, g, P, nti, ntk, jchunk, rho0);
            return hipGetLastError();
        }
    }
#define TDLB_SIM(M) \
    hipLaunchKernelGGL((k_simple<PR, L, M>), dim3(nb), dim3(bk), 0, s, a, b, mo, f, nu, fm, g, P, nkb, rho0)
    switch (mode) {
        case MODE_COLL: TDLB_SIM(MODE_COLL); break;
        case MODE_MOM: TDLB_SIM(MODE_MOM); break;
        case MODE_FUSED: TDLB_SIM(MODE_FUSED); break;
        case MODE_MOM_PULL: TDLB_SIM(MODE_MOM_PULL); break;
        default: return hipErrorInvalidValue;
    }
#undef TDLB_SIM
    return hipGetLastError();
}

template <typename PR, int L, int STREAM>
hipError_t force_level_t(int momsrc, int even, const PointsDev& p, int64_t begin, int64_t count, void* F,
                         const void* Q, const Grid& g, double alfa, double beta, float rho0, hipStream_t s) {
    using S = typename PR::S;
    using T = typename PR::T;
    const uint32_t nb = (uint32_t)((count + kBlock - 1) / kBlock);
    T* f = static_cast<T*>(F);
    const S* q = static_cast<const S*>(Q);
    const T a = (T)alfa, b = (T)beta;
#define TDLB_FL(M, E) \
    hipLaunchKernelGGL((k_force_level<PR, L, STREAM, M, E>), dim3(nb), dim3(kBlock), 0, s, p, begin, count, f, q, g, a, b, rho0)
    if (momsrc == MOMSRC_STORED) {
        if constexpr (PR::kShift) return hipErrorInvalidValue;
        else { if (even) TDLB_FL(MOMSRC_STORED, 1); else TDLB_FL(MOMSRC_STORED, 0); }
    }
    else if (momsrc == MOMSRC_ESO_POPS) { if (even) TDLB_FL(MOMSRC_ESO_POPS, 1); else TDLB_FL(MOMSRC_ESO_POPS, 0); }
    else { TDLB_FL(MOMSRC_PULL_POPS, 1); }
#undef TDLB_FL
    return hipGetLastError();
}

}  // namespace

// precision index -> traits: 0 fp32, 1 fp64, 2 fp16 storage, 3 bf16 storage,
// 4 e3m12 storage (the C ABI validates the index at tdlb_cu_create)
#define TDLB_WITH_PREC(prec, ...)                      \
    do {                                               \
        switch (prec) {                                \
            case 0: { using PR = PrecF32; __VA_ARGS__; } break;  \
            case 1: { using PR = PrecF64; __VA_ARGS__; } break;  \
            case 2: { using PR = PrecF16; __VA_ARGS__; } break;  \
            case 3: { using PR = PrecBF16; __VA_ARGS__; } break; \
            default: { using PR = PrecE3M12; __VA_ARGS__; } break; \
        }                                              \
    } while (0)

// ====================================================================== launchers
hipError_t launch_fill_q(int prec, int layout, void* Q, const Grid& g, double rho, hipStream_t s) {
    const uint32_t nb = grid_blocks((layout == LayoutIJKL ? g.size : g.qsize) * kQ);
    if (prec == 0) {
        if (layout == LayoutIJKL) hipLaunchKernelGGL((k_fill_q<float, LayoutIJKL>), dim3(nb), dim3(kBlock), 0, s, (float*)Q, g, (float)rho);
        else hipLaunchKernelGGL((k_fill_q<float, LayoutLIJK>), dim3(nb), dim3(kBlock), 0, s, (float*)Q, g, (float)rho);
    } else {
        if (layout == LayoutIJKL) hipLaunchKernelGGL((k_fill_q<double, LayoutIJKL>), dim3(nb), dim3(kBlock), 0, s, (double*)Q, g, rho);
        else hipLaunchKernelGGL((k_fill_q<double, LayoutLIJK>), dim3(nb), dim3(kBlock), 0, s, (double*)Q, g, rho);
    }
    return hipGetLastError();
}

namespace {
template <typename W>
hipError_t q_boundary_t(int dlayout, int blayout, bool to_boundary, const void* src, void* dst, const Grid& g,
                        hipStream_t s) {
    const uint32_t nb = grid_blocks(g.size * kQ);
    const W* a = static_cast<const W*>(src);
    W* b = static_cast<W*>(dst);
#define TDLB_QB(DL, BL)                                                                                   \
    do {                                                                                                  \
        if (to_boundary) hipLaunchKernelGGL((k_q_boundary<W, DL, BL, true>), dim3(nb), dim3(kBlock), 0, s, a, b, g); \
        else hipLaunchKernelGGL((k_q_boundary<W, DL, BL, false>), dim3(nb), dim3(kBlock), 0, s, a, b, g); \
    } while (0)
    if (dlayout == LayoutIJKL && blayout == LayoutIJKL) TDLB_QB(LayoutIJKL, LayoutIJKL);
    else if (dlayout == LayoutIJKL) TDLB_QB(LayoutIJKL, LayoutLIJK);
    else if (blayout == LayoutIJKL) TDLB_QB(LayoutLIJK, LayoutIJKL);
    else TDLB_QB(LayoutLIJK, LayoutLIJK);
#undef TDLB_QB
    return hipGetLastError();
}
}  // namespace

hipError_t launch_q_boundary(int esize, int dlayout, int blayout, int to_boundary, const void* src, void* dst,
                             const Grid& g, hipStream_t s) {
    if (esize == 4) return q_boundary_t<uint32_t>(dlayout, blayout, to_boundary != 0, src, dst, g, s);
    if (esize == 8) return q_boundary_t<uint64_t>(dlayout, blayout, to_boundary != 0, src, dst, g, s);
    return hipErrorInvalidValue;
}

// Speed of light of the stream-collide access pattern: one pass reads the
// value of every lane's cell from each of `planes` planes, then writes each
// back with its bits flipped by a run-time mask (so no load or store folds
// away; two passes restore the data).  Blocks of kRowBlock lanes, as the
// stream-collide rows.
constexpr int kProbePlanesMax = 20;
template <typename U, bool NT>
__global__ __launch_bounds__(kRowBlock) void k_probe_planes(U* __restrict__ base, uint64_t stride, uint64_t n,
                                                            int planes, U mask) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    U v[kProbePlanesMax];
#pragma unroll
    for (int p = 0; p < kProbePlanesMax; p++)
        if (p < planes) v[p] = load_hint<NT>(base + (uint64_t)p * stride + c);
#pragma unroll
    for (int p = 0; p < kProbePlanesMax; p++)
        if (p < planes) store_hint<NT>(base + (uint64_t)p * stride + c, (U)(v[p] ^ mask));
}

template <typename U>
hipError_t probe_t(void* base, uint64_t stride, uint64_t n, int planes, int nt, hipStream_t s) {
    const uint32_t nb = (uint32_t)((n + kRowBlock - 1) / kRowBlock);
    const U mask = (U)0x5555555555555555ull;
    if (nt) hipLaunchKernelGGL((k_probe_planes<U, true>), dim3(nb), dim3(kRowBlock), 0, s, (U*)base, stride, n, planes, mask);
    else hipLaunchKernelGGL((k_probe_planes<U, false>), dim3(nb), dim3(kRowBlock), 0, s, (U*)base, stride, n, planes, mask);
    return hipGetLastError();
}

hipError_t launch_probe_planes(int elem_bytes, void* base, uint64_t stride, uint64_t n, int planes, int nt,
                               hipStream_t s) {
    if (planes < 1 || planes > kProbePlanesMax || n == 0 || (n + kRowBlock - 1) / kRowBlock > 0x7fffffffull)
        return hipErrorInvalidValue;
    if (elem_bytes == 2) return probe_t<uint16_t>(base, stride, n, planes, nt, s);
    if (elem_bytes == 4) return probe_t<uint32_t>(base, stride, n, planes, nt, s);
    if (elem_bytes == 8) return probe_t<uint64_t>(base, stride, n, planes, nt, s);
    return hipErrorInvalidValue;
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    if (bytes % 4 == 0 && ((uintptr_t)p % 4) == 0)
        hipLaunchKernelGGL(k_zero_words, dim3(grid_blocks(bytes / 4)), dim3(kBlock), 0, s, (uint32_t*)p, (uint64_t)(bytes / 4));
    else
        hipLaunchKernelGGL(k_zero_bytes, dim3(grid_blocks(bytes)), dim3(kBlock), 0, s, (uint8_t*)p, (uint64_t)bytes);
    return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    if (bytes % 4 == 0 && ((uintptr_t)dst % 4) == 0 && ((uintptr_t)src % 4) == 0)
        hipLaunchKernelGGL(k_copy_words, dim3(grid_blocks(bytes / 4)), dim3(kBlock), 0, s, (uint32_t*)dst,
                           (const uint32_t*)src, (uint64_t)(bytes / 4));
    else
        hipLaunchKernelGGL(k_copy_bytes, dim3(grid_blocks(bytes)), dim3(kBlock), 0, s, (uint8_t*)dst,
                           (const uint8_t*)src, (uint64_t)bytes);
    return hipGetLastError();
}

hipError_t launch_esotwist(int prec, int layout, int mode, int even, void* Q, void* mout, const void* F, void* Nu,
                           const uint64_t* fmask, const Grid& g, const CollideArgs& c, float rho0, hipStream_t s) {
    hipError_t e = hipErrorInvalidValue;
    TDLB_WITH_PREC(prec, e = layout == LayoutIJKL
                                 ? esotwist_t<PR, LayoutIJKL>(mode, even, Q, mout, F, Nu, fmask, g, c, rho0, s)
                                 : esotwist_t<PR, LayoutLIJK>(mode, even, Q, mout, F, Nu, fmask, g, c, rho0, s));
    return e;
}

hipError_t launch_simple(int prec, int layout, int mode, const void* src, void* dst, void* mout, const void* F,
                         void* Nu, const uint64_t* fmask, const Grid& g, const CollideArgs& c, float rho0,
                         hipStream_t s) {
    hipError_t e = hipErrorInvalidValue;
    TDLB_WITH_PREC(prec, e = layout == LayoutIJKL
                                 ? simple_t<PR, LayoutIJKL>(mode, src, dst, mout, F, Nu, fmask, g, c, rho0, s)
                                 : simple_t<PR, LayoutLIJK>(mode, src, dst, mout, F, Nu, fmask, g, c, rho0, s));
    return e;
}

// launch KERNEL<W, layout> over NTHREADS grid-stride threads
#define TDLB_MOVE(KERNEL, W, NTHREADS, ...)                                                                     \
    do {                                                                                                        \
        const uint32_t nb_ = grid_blocks(NTHREADS);                                                             \
        if (layout == LayoutIJKL)                                                                               \
            hipLaunchKernelGGL((KERNEL<W, LayoutIJKL>), dim3(nb_), dim3(kBlock), 0, s, __VA_ARGS__);            \
        else                                                                                                    \
            hipLaunchKernelGGL((KERNEL<W, LayoutLIJK>), dim3(nb_), dim3(kBlock), 0, s, __VA_ARGS__);            \
    } while (0)

hipError_t launch_push_all(int esize, int layout, const void* src, void* dst, const Grid& g, hipStream_t s) {
    if (esize == 4) TDLB_MOVE(k_push_all, uint32_t, g.size, (const uint32_t*)src, (uint32_t*)dst, g);
    else TDLB_MOVE(k_push_all, uint64_t, g.size, (const uint64_t*)src, (uint64_t*)dst, g);
    return hipGetLastError();
}

hipError_t launch_pull_all(int esize, int layout, const void* src, void* dst, const Grid& g, hipStream_t s) {
    if (esize == 4) TDLB_MOVE(k_pull_all, uint32_t, g.size, (const uint32_t*)src, (uint32_t*)dst, g);
    else TDLB_MOVE(k_pull_all, uint64_t, g.size, (const uint64_t*)src, (uint64_t*)dst, g);
    return hipGetLastError();
}

hipError_t launch_bounce_back(int esize, int layout, void* Q, const Grid& g, hipStream_t s) {
    for (int face = 0; face < 6; face++) {
        const int axis = face >> 1;
        const uint64_t n = (uint64_t)(axis == 0 ? g.yg1 : g.xg1) * (axis == 2 ? g.yg1 : g.zg1);
        if (esize == 4) TDLB_MOVE(k_bb_face, uint32_t, n, (uint32_t*)Q, g, face);
        else TDLB_MOVE(k_bb_face, uint64_t, n, (uint64_t*)Q, g, face);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const uint64_t n = 12ull * (uint64_t)std::max(g.xg1, std::max(g.yg1, g.zg1));
    if (esize == 4) TDLB_MOVE(k_bb_edges, uint32_t, n, (uint32_t*)Q, g);
    else TDLB_MOVE(k_bb_edges, uint64_t, n, (uint64_t*)Q, g);
    return hipGetLastError();
}

hipError_t launch_ghost_fill(int esize, int layout, const void* src, void* dst, const Grid& g, hipStream_t s) {
    if (esize == 4)
        TDLB_MOVE(k_ghost_fill, RawBits<uint32_t>, ghost_count(g), (const uint32_t*)src, (uint32_t*)dst, g, 0.0f);
    else
        TDLB_MOVE(k_ghost_fill, RawBits<uint64_t>, ghost_count(g), (const uint64_t*)src, (uint64_t*)dst, g, 0.0f);
    return hipGetLastError();
}

hipError_t launch_ghost_fill_decode(int prec, int layout, const void* src, void* dst, const Grid& g, float rho0,
                                    hipStream_t s) {
    if (prec == 2)
        TDLB_MOVE(k_ghost_fill, PrecF16, ghost_count(g), (const PrecF16::S*)src, (float*)dst, g, rho0);
    else if (prec == 3)
        TDLB_MOVE(k_ghost_fill, PrecBF16, ghost_count(g), (const PrecBF16::S*)src, (float*)dst, g, rho0);
    else if (prec == 4)
        TDLB_MOVE(k_ghost_fill, PrecE3M12, ghost_count(g), (const PrecE3M12::S*)src, (float*)dst, g, rho0);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_ghost_copy(int esize, int layout, const void* src, void* dst, const Grid& g, hipStream_t s) {
    if (esize == 4) TDLB_MOVE(k_ghost_copy, uint32_t, ghost_count(g), (const uint32_t*)src, (uint32_t*)dst, g);
    else TDLB_MOVE(k_ghost_copy, uint64_t, ghost_count(g), (const uint64_t*)src, (uint64_t*)dst, g);
    return hipGetLastError();
}
#undef TDLB_MOVE

hipError_t launch_mask_from_f(int prec, const void* F, uint64_t* mask, uint64_t* gmask, const Grid& g,
                              hipStream_t s) {
    const uint64_t words = (uint64_t)g.xg * g.yg * g.mask_w;
    if (prec != 1) hipLaunchKernelGGL(k_mask_from_f<float>, dim3(grid_blocks(words)), dim3(kBlock), 0, s, (const float*)F, mask, gmask, g);
    else hipLaunchKernelGGL(k_mask_from_f<double>, dim3(grid_blocks(words)), dim3(kBlock), 0, s, (const double*)F, mask, gmask, g);
    return hipGetLastError();
}

hipError_t launch_masked_f(int prec, const void* F, const uint64_t* mask, void* out, const Grid& g, hipStream_t s) {
    if (prec != 1) hipLaunchKernelGGL(k_masked_f<float>, dim3(grid_blocks(g.size)), dim3(kBlock), 0, s, (const float*)F, mask, (float*)out, g);
    else hipLaunchKernelGGL(k_masked_f<double>, dim3(grid_blocks(g.size)), dim3(kBlock), 0, s, (const double*)F, mask, (double*)out, g);
    return hipGetLastError();
}

hipError_t launch_zero_interior_o(uint8_t* O, const Grid& g, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_interior_o, dim3(grid_blocks(g.size)), dim3(kBlock), 0, s, O, g);
    return hipGetLastError();
}

hipError_t launch_force_prepare(int prec, const PointsDev& p, int64_t begin, int64_t count, void* F,
                                const uint64_t* prev, uint64_t* next, const Grid& g, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const uint32_t nb = (uint32_t)((count + kBlock - 1) / kBlock);
    if (prec != 1) hipLaunchKernelGGL(k_force_prepare<float>, dim3(nb), dim3(kBlock), 0, s, p, begin, count, (float*)F, prev, next, g);
    else hipLaunchKernelGGL(k_force_prepare<double>, dim3(nb), dim3(kBlock), 0, s, p, begin, count, (double*)F, prev, next, g);
    return hipGetLastError();
}

hipError_t launch_force_level(int prec, int layout, int streaming, int momsrc, int even, const PointsDev& p,
                              int64_t begin, int64_t count, void* F, const void* Q, const Grid& g, double alfa,
                              double beta, float rho0, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipError_t e = hipErrorInvalidValue;
    TDLB_WITH_PREC(prec, {
        if (layout == LayoutIJKL)
            e = streaming == StreamEsotwist
                    ? force_level_t<PR, LayoutIJKL, StreamEsotwist>(momsrc, even, p, begin, count, F, Q, g, alfa, beta, rho0, s)
                    : force_level_t<PR, LayoutIJKL, StreamSimple>(momsrc, even, p, begin, count, F, Q, g, alfa, beta, rho0, s);
        else
            e = streaming == StreamEsotwist
                    ? force_level_t<PR, LayoutLIJK, StreamEsotwist>(momsrc, even, p, begin, count, F, Q, g, alfa, beta, rho0, s)
                    : force_level_t<PR, LayoutLIJK, StreamSimple>(momsrc, even, p, begin, count, F, Q, g, alfa, beta, rho0, s);
    });
    return e;
}

namespace {
template <typename PR, int L, int STREAM>
hipError_t ib_moments_t(int momsrc, int even, const uint32_t* cells, const PointsDev* p, int64_t begin, int64_t n,
                        void* cache, const void* Q, const Grid& g, float rho0, hipStream_t s) {
    using T = typename PR::T;
    using S = typename PR::S;
    Mom4<T>* c = static_cast<Mom4<T>*>(cache);
    const S* q = static_cast<const S*>(Q);
    const int64_t threads = cells ? n : 9 * n;
    const uint32_t nb = (uint32_t)((threads + kBlock - 1) / kBlock);
#define TDLB_IBM(M, E)                                                                                          \
    do {                                                                                                        \
        if (cells) hipLaunchKernelGGL((k_ib_moments_cells<PR, L, STREAM, M, E>), dim3(nb), dim3(kBlock), 0, s, cells, n, c, q, g, rho0); \
        else hipLaunchKernelGGL((k_ib_moments_points<PR, L, STREAM, M, E>), dim3(nb), dim3(kBlock), 0, s, *p, begin, n, c, q, g, rho0); \
    } while (0)
    if (momsrc == MOMSRC_STORED) {
        if constexpr (PR::kShift) return hipErrorInvalidValue;
        else { if (even) TDLB_IBM(MOMSRC_STORED, 1); else TDLB_IBM(MOMSRC_STORED, 0); }
    }
    else if (momsrc == MOMSRC_ESO_POPS) { if (even) TDLB_IBM(MOMSRC_ESO_POPS, 1); else TDLB_IBM(MOMSRC_ESO_POPS, 0); }
    else TDLB_IBM(MOMSRC_PULL_POPS, 1);
#undef TDLB_IBM
    return hipGetLastError();
}

template <typename PR>
hipError_t ib_moments_prec(int layout, int streaming, int momsrc, int even, const uint32_t* cells, const PointsDev* p,
                           int64_t begin, int64_t n, void* cache, const void* Q, const Grid& g, float rho0,
                           hipStream_t s) {
    if (layout == LayoutIJKL)
        return streaming == StreamEsotwist
                   ? ib_moments_t<PR, LayoutIJKL, StreamEsotwist>(momsrc, even, cells, p, begin, n, cache, Q, g, rho0, s)
                   : ib_moments_t<PR, LayoutIJKL, StreamSimple>(momsrc, even, cells, p, begin, n, cache, Q, g, rho0, s);
    return streaming == StreamEsotwist
               ? ib_moments_t<PR, LayoutLIJK, StreamEsotwist>(momsrc, even, cells, p, begin, n, cache, Q, g, rho0, s)
               : ib_moments_t<PR, LayoutLIJK, StreamSimple>(momsrc, even, cells, p, begin, n, cache, Q, g, rho0, s);
}
}  // namespace

hipError_t launch_ib_moments_cells(int prec, int layout, int streaming, int momsrc, int even, const uint32_t* cells,
                                   int64_t n, void* cache, const void* Q, const Grid& g, float rho0, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipError_t e = hipErrorInvalidValue;
    TDLB_WITH_PREC(prec, e = ib_moments_prec<PR>(layout, streaming, momsrc, even, cells, nullptr, 0, n, cache, Q, g,
                                                 rho0, s));
    return e;
}

hipError_t launch_ib_moments_points(int prec, int layout, int streaming, int momsrc, int even, const PointsDev& p,
                                    int64_t begin, int64_t count, void* cache, const void* Q, const Grid& g,
                                    float rho0, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipError_t e = hipErrorInvalidValue;
    TDLB_WITH_PREC(prec, e = ib_moments_prec<PR>(layout, streaming, momsrc, even, nullptr, &p, begin, count, cache, Q,
                                                 g, rho0, s));
    return e;
}

hipError_t launch_force_level_cached(int prec, const PointsDev& p, int64_t begin, int64_t count,
                                     const int32_t* sidx, int64_t cache_base, const void* cache, void* F,
                                     const Grid& g, double alfa, double beta, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    constexpr int bs = 64;  // one wavefront per block: levels are small, spread them over the CUs
    const uint32_t nb = group_blocks(count, bs);
    TDLB_WITH_PREC(prec, {
        using T = typename PR::T;
        using D = typename PR::D;
        const auto* c = static_cast<const Mom4<T>*>(cache);
        if (sidx)
            hipLaunchKernelGGL((k_force_level_cached<T, D, true>), dim3(nb), dim3(bs), 0, s, p, begin, count, sidx,
                               cache_base, c, (T*)F, g, (T)alfa, (T)beta);
        else
            hipLaunchKernelGGL((k_force_level_cached<T, D, false>), dim3(nb), dim3(bs), 0, s, p, begin, count, sidx,
                               cache_base, c, (T*)F, g, (T)alfa, (T)beta);
    });
    return hipGetLastError();
}

namespace {
// fn(L, STREAM, MOMSRC, EVEN) with compile-time values (std::integral_constant)
template <int V>
using IC = std::integral_constant<int, V>;

template <typename PR, int L, int STREAM, typename Fn>
hipError_t with_source(int momsrc, int even, Fn&& fn) {
    if (momsrc == MOMSRC_STORED) {
        if constexpr (PR::kShift) return hipErrorInvalidValue;
        else return even ? fn(IC<L>{}, IC<STREAM>{}, IC<MOMSRC_STORED>{}, IC<1>{})
                         : fn(IC<L>{}, IC<STREAM>{}, IC<MOMSRC_STORED>{}, IC<0>{});
    }
    if (momsrc == MOMSRC_ESO_POPS) {
        if constexpr (STREAM != StreamEsotwist) return hipErrorInvalidValue;
        else return even ? fn(IC<L>{}, IC<STREAM>{}, IC<MOMSRC_ESO_POPS>{}, IC<1>{})
                         : fn(IC<L>{}, IC<STREAM>{}, IC<MOMSRC_ESO_POPS>{}, IC<0>{});
    }
    if constexpr (STREAM != StreamSimple) return hipErrorInvalidValue;
    else return fn(IC<L>{}, IC<STREAM>{}, IC<MOMSRC_PULL_POPS>{}, IC<1>{});
}

template <typename PR, typename Fn>
hipError_t with_cell_source(int layout, int streaming, int momsrc, int even, Fn&& fn) {
    if (layout == LayoutIJKL)
        return streaming == StreamEsotwist ? with_source<PR, LayoutIJKL, StreamEsotwist>(momsrc, even, fn)
                                           : with_source<PR, LayoutIJKL, StreamSimple>(momsrc, even, fn);
    return streaming == StreamEsotwist ? with_source<PR, LayoutLIJK, StreamEsotwist>(momsrc, even, fn)
                                       : with_source<PR, LayoutLIJK, StreamSimple>(momsrc, even, fn);
}
}  // namespace

hipError_t launch_force_planes(int prec, int layout, int streaming, int momsrc, int even, const ForceGeom& fg,
                               const void* Q, void* F, const uint64_t* prev, uint64_t* next, const Grid& g,
                               double alfa, double beta, float rho0, hipStream_t s) {
    if (fg.n_blocks <= 0) return hipSuccess;
    hipError_t e = hipErrorInvalidValue;
    TDLB_WITH_PREC(prec, {
        using S = typename PR::S;
        using T = typename PR::T;
        e = with_cell_source<PR>(layout, streaming, momsrc, even, [&](auto Lc, auto Sc, auto Mc, auto Ec) {
            hipLaunchKernelGGL((k_force_planes<PR, decltype(Lc)::value, decltype(Sc)::value, decltype(Mc)::value,
                                               decltype(Ec)::value>),
                               dim3((uint32_t)fg.n_blocks), dim3(kPlaneBlock), 0, s, fg, (const S*)Q, (T*)F, prev,
                               next, g, (T)alfa, (T)beta, rho0);
            return hipGetLastError();
        });
    });
    return e;
}

bool line_rows_bytes(const Grid& g, size_t ssize) {
    return ssize == 4 ? line_rows<float>(g) : line_rows<_Float16>(g);
}

// the pipe's row blocks take at most 256 lanes: 4 waves, so that several
// workgroups share a CU at the register count of the stream-collide code
constexpr int kPipeLanes = 256;
uint32_t row_lanes(const Grid& g) {
    uint32_t nkb, nb, bk;
    row_launch(g, nkb, nb, bk, kPipeLanes);
    return bk;
}

hipError_t launch_eso_pipe(int prec, int even, int hint, const PipeArgs& a, const Grid& g, const CollideArgs& c,
                           float rho0, double alfa, double beta, uint32_t blocks, uint32_t lanes, hipStream_t s) {
    uint32_t nkb, nb, bk;
    row_launch(g, nkb, nb, bk, kPipeLanes);
    if (lanes != bk) return hipErrorInvalidValue;
    auto go = [&](auto pr) -> hipError_t {
        using PR = decltype(pr);
        using S = typename PR::S;
        const auto P = to_params<float, typename PR::D>(c);
#define TDLB_PIPE(E, H)                                                                                         \
    hipLaunchKernelGGL((k_eso_pipe<PR, E, H>), dim3(blocks), dim3(bk), 0, s, a, g, P, nkb, rho0, (float)alfa,    \
                       (float)beta)
        if (blocks == 0) {
            // persistent grid: every workgroup resident at once
            int per_cu = 0, dev = 0;
            hipDeviceProp_t prop;
            if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_eso_pipe<PR, 1, true>, (int)bk, 0) !=
                    hipSuccess)
                return hipErrorInvalidValue;
            // at least 8 per CU: workgroups beyond the resident ones find the queues drained
            blocks = (uint32_t)std::max(8, per_cu) * (uint32_t)prop.multiProcessorCount;
        }
        if (even) { if (hint) TDLB_PIPE(1, true); else TDLB_PIPE(1, false); }
        else { if (hint) TDLB_PIPE(0, true); else TDLB_PIPE(0, false); }
#undef TDLB_PIPE
        hipError_t e = hipGetLastError();
        if (e != hipSuccess || a.fg.n_blocks <= 0) return e;
        // the levels read the populations of the next step's collision
        const S* q = static_cast<const S*>(a.Q);
        if (even)
            hipLaunchKernelGGL((k_force_levels<PR, 0>), dim3((uint32_t)a.fg.n_blocks), dim3(kLevelBlock), 0, s, a.fg,
                               q, static_cast<float*>(a.F), a.prev, a.next, g, (float)alfa, (float)beta, rho0);
        else
            hipLaunchKernelGGL((k_force_levels<PR, 1>), dim3((uint32_t)a.fg.n_blocks), dim3(kLevelBlock), 0, s, a.fg,
                               q, static_cast<float*>(a.F), a.prev, a.next, g, (float)alfa, (float)beta, rho0);
        return hipGetLastError();
    };
    switch (prec) {
        case 0: return go(PrecF32{});
        case 2: return go(PrecF16{});
        case 3: return go(PrecBF16{});
        case 4: return go(PrecE3M12{});
        default: return hipErrorInvalidValue;
    }
}

int pipe_xcc_check(hipStream_t s) {
    constexpr int nb = 64 * kXcds;
    uint32_t* d = nullptr;
    if (hipMalloc((void**)&d, nb * sizeof(uint32_t)) != hipSuccess) return 0;
    hipLaunchKernelGGL(k_xcc_ids, dim3(nb), dim3(64), 0, s, d);
    uint32_t h[nb];
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s) == hipSuccess &&
                    hipStreamSynchronize(s) == hipSuccess;
    (void)hipFree(d);
    if (!ok) return 0;
    int seen[kXcds] = {0};
    for (int b = 0; b < nb; b++) {
        if (h[b] >= (uint32_t)kXcds) return 0;
        seen[h[b]]++;
    }
    for (int x = 0; x < kXcds; x++)
        if (seen[x] == 0) return 0;
    return 1;
}

hipError_t launch_reduce(int prec, int layout, int streaming, int momsrc, int even, const void* Q, const void* F,
                         const uint64_t* fmask, const Grid& g, const ReduceArgs& ra, float rho0, double* partial,
                         double* out, hipStream_t s) {
    hipError_t e = hipErrorInvalidValue;
    TDLB_WITH_PREC(prec, {
        using S = typename PR::S;
        using T = typename PR::T;
        e = with_cell_source<PR>(layout, streaming, momsrc, even, [&](auto Lc, auto Sc, auto Mc, auto Ec) {
            hipLaunchKernelGGL((k_reduce_cells<PR, decltype(Lc)::value, decltype(Sc)::value, decltype(Mc)::value,
                                               decltype(Ec)::value>),
                               dim3(kReduceBlocks), dim3(kBlock), 0, s, (const S*)Q, (const T*)F, fmask, g, ra, rho0,
                               partial);
            return hipGetLastError();
        });
    });
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, s, partial, kReduceBlocks, out);
    return hipGetLastError();
}

hipError_t launch_macroscopic(int prec, int layout, int streaming, int momsrc, int even, const void* Q,
                              const void* F, const uint64_t* fmask, const Grid& g, float rho0, void* out,
                              hipStream_t s) {
    const uint32_t nb = grid_blocks((uint64_t)g.xg1 * g.yg1 * g.zg1);
    hipError_t e = hipErrorInvalidValue;
    TDLB_WITH_PREC(prec, {
        using S = typename PR::S;
        using T = typename PR::T;
        e = with_cell_source<PR>(layout, streaming, momsrc, even, [&](auto Lc, auto Sc, auto Mc, auto Ec) {
            hipLaunchKernelGGL((k_macroscopic<PR, decltype(Lc)::value, decltype(Sc)::value, decltype(Mc)::value,
                                              decltype(Ec)::value>),
                               dim3(nb), dim3(kBlock), 0, s, (const S*)Q, (const T*)F, fmask, g, rho0, (T*)out);
            return hipGetLastError();
        });
    });
    return e;
}

hipError_t launch_regrid(int esize, int layout, const void* oldQ, void* newQ, int nu_esize, const void* oldNu,
                         void* newNu, const Grid& og, const Grid& ng, hipStream_t s) {
    const uint32_t nb = grid_blocks(ng.size);
    if (esize == 4) {
        if (layout == LayoutIJKL)
            hipLaunchKernelGGL((k_regrid_q<uint32_t, LayoutIJKL>), dim3(nb), dim3(kBlock), 0, s, (const uint32_t*)oldQ, (uint32_t*)newQ, og, ng);
        else
            hipLaunchKernelGGL((k_regrid_q<uint32_t, LayoutLIJK>), dim3(nb), dim3(kBlock), 0, s, (const uint32_t*)oldQ, (uint32_t*)newQ, og, ng);
    } else if (esize == 8) {
        if (layout == LayoutIJKL)
            hipLaunchKernelGGL((k_regrid_q<uint64_t, LayoutIJKL>), dim3(nb), dim3(kBlock), 0, s, (const uint64_t*)oldQ, (uint64_t*)newQ, og, ng);
        else
            hipLaunchKernelGGL((k_regrid_q<uint64_t, LayoutLIJK>), dim3(nb), dim3(kBlock), 0, s, (const uint64_t*)oldQ, (uint64_t*)newQ, og, ng);
    } else {
        return hipErrorInvalidValue;
    }
    if (nu_esize == 2)
        hipLaunchKernelGGL(k_regrid_cells<uint16_t>, dim3(nb), dim3(kBlock), 0, s, (const uint16_t*)oldNu, (uint16_t*)newNu, og, ng);
    else if (nu_esize == 4)
        hipLaunchKernelGGL(k_regrid_cells<uint32_t>, dim3(nb), dim3(kBlock), 0, s, (const uint32_t*)oldNu, (uint32_t*)newNu, og, ng);
    else if (nu_esize == 8)
        hipLaunchKernelGGL(k_regrid_cells<uint64_t>, dim3(nb), dim3(kBlock), 0, s, (const uint64_t*)oldNu, (uint64_t*)newNu, og, ng);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

namespace {
// Nu between the 16-bit storage of prec 2/3/4 and the float array of the boundary
template <typename PR, bool TO_FLOAT>
__global__ void k_nu_convert(const void* __restrict__ src, void* __restrict__ dst, uint64_t n) {
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n; a += (uint64_t)gridDim.x * blockDim.x) {
        if (TO_FLOAT) static_cast<float*>(dst)[a] = PR::dec_nu(static_cast<const typename PR::N*>(src)[a]);
        else static_cast<typename PR::N*>(dst)[a] = PR::enc_nu(static_cast<const float*>(src)[a]);
    }
}
}  // namespace

hipError_t launch_nu_convert(int prec, int to_float, const void* src, void* dst, uint64_t n, hipStream_t s) {
    const uint32_t nb = grid_blocks(n);
    if (prec == 2) {
        if (to_float) hipLaunchKernelGGL((k_nu_convert<PrecF16, true>), dim3(nb), dim3(kBlock), 0, s, src, dst, n);
        else hipLaunchKernelGGL((k_nu_convert<PrecF16, false>), dim3(nb), dim3(kBlock), 0, s, src, dst, n);
    } else if (prec == 3) {
        if (to_float) hipLaunchKernelGGL((k_nu_convert<PrecBF16, true>), dim3(nb), dim3(kBlock), 0, s, src, dst, n);
        else hipLaunchKernelGGL((k_nu_convert<PrecBF16, false>), dim3(nb), dim3(kBlock), 0, s, src, dst, n);
    } else if (prec == 4) {
        if (to_float) hipLaunchKernelGGL((k_nu_convert<PrecE3M12, true>), dim3(nb), dim3(kBlock), 0, s, src, dst, n);
        else hipLaunchKernelGGL((k_nu_convert<PrecE3M12, false>), dim3(nb), dim3(kBlock), 0, s, src, dst, n);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_plane_xz(int prec, int layout, const void* Q, const void* F, const uint64_t* fmask,
                           const uint8_t* exclude, const Grid& g, int cut, int64_t i0, int64_t k0, PlaneRec4* q,
                           PlaneRec3* f, uint8_t* flags, hipStream_t s) {
    const uint32_t nb = (uint32_t)(((int64_t)g.xg1 * g.zg1 + kBlock - 1) / kBlock);
#define TDLB_PLANE(T, L)                                                                                       \
    hipLaunchKernelGGL((k_plane_xz<T, L>), dim3(nb), dim3(kBlock), 0, s, (const T*)Q, (const T*)F, fmask, exclude, \
                       g, cut, i0, k0, q, f, flags)
    if (prec == 1) {
        if (layout == LayoutIJKL) TDLB_PLANE(double, LayoutIJKL);
        else TDLB_PLANE(double, LayoutLIJK);
    } else {
        if (layout == LayoutIJKL) TDLB_PLANE(float, LayoutIJKL);
        else TDLB_PLANE(float, LayoutLIJK);
    }
#undef TDLB_PLANE
    return hipGetLastError();
}

hipError_t launch_vorticity(int prec, int layout, int streaming, int momsrc, int even, const void* Q, const Grid& g,
                            int axis, int cut, const uint8_t* exclude, void* out, float rho0, hipStream_t s) {
    const int64_t cells = (int64_t)g.xg1 * (axis == 0 ? g.zg1 : g.yg1);
    const uint32_t nb = (uint32_t)((cells + kBlock - 1) / kBlock);
    hipError_t e = hipErrorInvalidValue;
    TDLB_WITH_PREC(prec, {
        using S = typename PR::S;
        using T = typename PR::T;
        e = with_cell_source<PR>(layout, streaming, momsrc, even, [&](auto Lc, auto Sc, auto Mc, auto Ec) {
            hipLaunchKernelGGL((k_vorticity<PR, decltype(Lc)::value, decltype(Sc)::value, decltype(Mc)::value,
                                            decltype(Ec)::value>),
                               dim3(nb), dim3(kBlock), 0, s, (const S*)Q, g, axis, cut, exclude, (T*)out, rho0);
            return hipGetLastError();
        });
    });
    return e;
}

hipError_t launch_gen_blades(int prec, const BladeLayout& bl, const BladeStep& st, const PointsDev& out,
                             int64_t offset, hipStream_t s) {
    const int64_t count = (int64_t)bl.n_blades * bl.n_planes * bl.n_radial;
    if (count <= 0) return hipSuccess;
    const uint32_t nb = (uint32_t)((count + kBlock - 1) / kBlock);
    if (prec != 1) hipLaunchKernelGGL(k_gen_blades<float>, dim3(nb), dim3(kBlock), 0, s, bl, st, out, offset, count);
    else hipLaunchKernelGGL(k_gen_blades<double>, dim3(nb), dim3(kBlock), 0, s, bl, st, out, offset, count);
    return hipGetLastError();
}

hipError_t launch_solid_mask(double cx, double cz, double radius, uint8_t* out, const Grid& g, int ghost,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_solid_mask, dim3(grid_blocks(g.size)), dim3(kBlock), 0, s, cx, cz, radius, out, g, ghost);
    return hipGetLastError();
}

// ====================================================================== batches
// Each kernel below runs the per-member body of the single-handle kernel above
// (same device functions), so a member advances exactly as it would alone.
namespace {

// largest m in [0, n) with off[m] <= x (off non-decreasing, off[0] <= x)
__device__ __forceinline__ int find_slot(const int64_t* __restrict__ off, int n, int64_t x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// interior cell of thread threadIdx.x in block b of a member's fused launch
__device__ __forceinline__ bool batch_cell(const Grid& g, int tile, uint32_t b, int& i, int& j, int& k) {
    if (tile == TILE_ROW64) {
        const uint32_t segs = (uint32_t)(g.zg1 + 63) >> 6;
        const uint32_t w = b * (blockDim.x >> 6) + (threadIdx.x >> 6);
        const uint32_t rows = (uint32_t)g.xg1 * (uint32_t)g.yg1;
        if (w >= rows * segs) return false;
        const uint32_t row = w / segs, seg = w - row * segs;
        const uint32_t ii = row / (uint32_t)g.yg1;
        i = (int)ii + 1;
        j = (int)(row - ii * (uint32_t)g.yg1) + 1;
        k = (int)(seg * 64 + (threadIdx.x & 63)) + 1;
        return k <= g.zg1;
    }
    if (tile == TILE_FLAT) {
        const uint32_t plane = (uint32_t)g.yg1 * (uint32_t)g.zg1;
        const uint64_t t = (uint64_t)b * blockDim.x + threadIdx.x;
        if (t >= (uint64_t)g.xg1 * plane) return false;
        const uint32_t ip = (uint32_t)(t / plane), r = (uint32_t)(t - (uint64_t)ip * plane);
        i = (int)ip + 1;
        j = (int)(r / (uint32_t)g.zg1) + 1;
        k = (int)(r - (uint32_t)(j - 1) * (uint32_t)g.zg1) + 1;
        return true;
    }
    const uint32_t plane = (uint32_t)g.yg * (uint32_t)g.zg;
    const uint32_t bpp = (plane + blockDim.x - 1) / blockDim.x;
    const uint32_t ip = b / bpp;
    const uint32_t local = (b - ip * bpp) * blockDim.x + threadIdx.x;
    if (ip >= (uint32_t)g.xg1 || local >= plane) return false;
    i = (int)ip + 1;
    j = (int)(local / (uint32_t)g.zg);
    k = (int)(local - (uint32_t)j * (uint32_t)g.zg);
    return j >= 1 && j <= g.yg1 && k >= 1 && k <= g.zg1;
}

template <typename PR, int L, int STREAM>
__global__ __launch_bounds__(kRowBlock) void k_batch_fused(const MemberDev* __restrict__ mem,
                                                        const int64_t* __restrict__ off, int n,
                                                        const int32_t* __restrict__ bm) {
    static_assert(std::is_same<typename PR::T, float>::value, "batches compute in float");
    using S = typename PR::S;
    const uint32_t blk = layout_block<L>();
    // one load from the block table instead of a binary search of dependent loads
    const int m = bm ? bm[blk] : find_slot(off, n, blk);
    const MemberDev& md = mem[m];
    if (!md.active) return;
    int i, j, k;
    if (!batch_cell(md.g, md.tile, (uint32_t)(blk - off[m]), i, j, k)) return;
    const uint64_t* fm = md.fmask[md.fm];
    if constexpr (STREAM == StreamEsotwist) {
        S* Q = static_cast<S*>(md.Q[0]);
        auto* nu = static_cast<typename PR::N*>(md.Nu);
        if (md.even) esotwist_cell<PR, L, MODE_FUSED, 1>(Q, nullptr, md.F, nu, fm, md.g, md.P, md.rho0, i, j, k);
        else esotwist_cell<PR, L, MODE_FUSED, 0>(Q, nullptr, md.F, nu, fm, md.g, md.P, md.rho0, i, j, k);
    } else {
        simple_cell<PR, L, MODE_FUSED>(static_cast<const S*>(md.Q[md.cur]), static_cast<S*>(md.Q[1 - md.cur]),
                                       nullptr, md.F, static_cast<typename PR::N*>(md.Nu), fm, md.g, md.P,
                                       md.rho0, i, j, k);
    }
}

__global__ void k_batch_blades(const MemberDev* __restrict__ mem, const int64_t* __restrict__ off, int n,
                               int64_t total) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total || t >= off[n]) return;
    const int m = find_slot(off, n, t);
    const MemberDev& md = mem[m];
    if (!md.active) return;
    blade_gen_point<float>(md.fg.blade, md.steps[md.done], md.fg.pts, md.fg.n_static, t - off[m]);
}

// the new occupancy mask starts from the static stencil bits
__global__ void k_batch_mask_init(const MemberDev* __restrict__ mem, const int64_t* __restrict__ off, int n,
                                  int64_t total) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total || t >= off[n]) return;
    const int m = find_slot(off, n, t);
    const MemberDev& md = mem[m];
    if (!md.active) return;
    md.fmask[1 - md.fm][t - off[m]] = md.static_base[t - off[m]];
}

// the forcing planes of every member: one workgroup per (member, plane) task,
// density and momentum from the populations the fused kernel just wrote
template <typename PR, int L, int STREAM>
__global__ __launch_bounds__(kPlaneBlock) void k_batch_force_planes(const MemberDev* __restrict__ mem,
                                                                    const PlaneTask* __restrict__ tasks,
                                                                    int64_t lds_cells) {
    extern __shared__ float4 batch_lds[];  // lds_cells force values (3 T each), launch-sized
    const PlaneTask tk = tasks[blockIdx.x];
    const MemberDev& md = mem[tk.member];
    if (!md.active) return;
    using S = typename PR::S;
    using T = typename PR::T;
    T* lds = lds_cells > 0 ? reinterpret_cast<T*>(batch_lds) : nullptr;
    const PlaneBlock pb = md.fg.blocks[tk.block];
    if (pb.plane < 0) return;
    const PlaneDesc& pd = md.fg.planes[pb.plane];
    const uint64_t* prev = md.fmask[md.fm];
    uint64_t* next = md.fmask[1 - md.fm];
    if constexpr (STREAM == StreamEsotwist) {
        const S* Q = static_cast<const S*>(md.Q[0]);
        if (md.even)
            force_plane<PR, L, STREAM, MOMSRC_ESO_POPS, 0>(md.fg, pd, pb.rank, Q, md.F, prev, next, md.g, md.alfa, md.beta,
                                                            md.rho0, lds, lds_cells);
        else
            force_plane<PR, L, STREAM, MOMSRC_ESO_POPS, 1>(md.fg, pd, pb.rank, Q, md.F, prev, next, md.g, md.alfa, md.beta,
                                                            md.rho0, lds, lds_cells);
    } else {
        const S* Q = static_cast<const S*>(md.Q[1 - md.cur]);
        force_plane<PR, L, STREAM, MOMSRC_PULL_POPS, 1>(md.fg, pd, pb.rank, Q, md.F, prev, next, md.g, md.alfa, md.beta,
                                                         md.rho0, lds, lds_cells);
    }
}

template <int STREAM>
__global__ void k_batch_advance(MemberDev* __restrict__ mem, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    MemberDev& md = mem[m];
    if (!md.active) return;
    md.done += 1;
    if (STREAM == StreamEsotwist) md.even ^= 1;
    else md.cur ^= 1;
    md.fm ^= 1;
    if (md.done >= md.n_steps) md.active = 0;
}

inline uint32_t nblk(int64_t n, int bs) { return (uint32_t)((n + bs - 1) / bs); }

template <typename PR, int L, int STREAM>
hipError_t batch_fused_t(const BatchPlan& p, hipStream_t s) {
    if (p.tot_coll <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_batch_fused<PR, L, STREAM>), dim3((uint32_t)p.tot_coll), dim3(p.bs), 0, s, p.mem,
                       p.off_coll, p.n_members, p.coll_member);
    return hipGetLastError();
}

template <typename PR, int L, int STREAM>
hipError_t batch_step_t(const BatchPlan& p, hipStream_t s) {
    const int n = p.n_members;
    if (p.bs != 256 && p.bs != 512) return hipErrorInvalidValue;
    if (p.tot_rot > 0)
        hipLaunchKernelGGL(k_batch_blades, dim3(nblk(p.tot_rot, kBlock)), dim3(kBlock), 0, s, p.mem, p.off_rot, n,
                           p.tot_rot);
    hipError_t e = batch_fused_t<PR, L, STREAM>(p, s);
    if (e != hipSuccess) return e;
    if (p.tot_words > 0)
        hipLaunchKernelGGL(k_batch_mask_init, dim3(nblk(p.tot_words, kBlock)), dim3(kBlock), 0, s, p.mem,
                           p.off_words, n, p.tot_words);
    if (p.n_tasks > 0)
        hipLaunchKernelGGL((k_batch_force_planes<PR, L, STREAM>), dim3((uint32_t)p.n_tasks),
                           dim3(p.force_bs > 0 ? p.force_bs : kPlaneBlock),
                           (uint32_t)(p.force_lds_cells * 3 * sizeof(typename PR::T)), s, p.mem, p.tasks,
                           p.force_lds_cells);
    hipLaunchKernelGGL(k_batch_advance<STREAM>, dim3(nblk(n, 64)), dim3(64), 0, s, p.mem_rw, n);
    return hipGetLastError();
}

template <typename Fn>
hipError_t with_batch_prec(int prec, int layout, int streaming, Fn&& fn) {
    auto by_layout = [&](auto pr) -> hipError_t {
        using PR = decltype(pr);
        if (layout == LayoutIJKL)
            return streaming == StreamEsotwist ? fn(PR{}, IC<LayoutIJKL>{}, IC<StreamEsotwist>{})
                                               : fn(PR{}, IC<LayoutIJKL>{}, IC<StreamSimple>{});
        return streaming == StreamEsotwist ? fn(PR{}, IC<LayoutLIJK>{}, IC<StreamEsotwist>{})
                                           : fn(PR{}, IC<LayoutLIJK>{}, IC<StreamSimple>{});
    };
    switch (prec) {
        case 0: return by_layout(PrecF32{});
        case 2: return by_layout(PrecF16{});
        case 3: return by_layout(PrecBF16{});
        case 4: return by_layout(PrecE3M12{});
        default: return hipErrorInvalidValue;
    }
}
}  // namespace

int64_t batch_coll_blocks(const Grid& g, int tile, int bs) {
    if (tile == TILE_ROW64) {
        const int64_t segs = (g.zg1 + 63) / 64;
        const int64_t waves = (int64_t)g.xg1 * g.yg1 * segs;
        return (waves + bs / 64 - 1) / (bs / 64);
    }
    if (tile == TILE_FLAT) return ((int64_t)g.xg1 * g.yg1 * g.zg1 + bs - 1) / bs;
    const int64_t plane = (int64_t)g.yg * g.zg;
    return (int64_t)g.xg1 * ((plane + bs - 1) / bs);
}

hipError_t launch_batch_step(int prec, int layout, int streaming, const BatchPlan& plan, int steps, hipStream_t s) {
    return with_batch_prec(prec, layout, streaming, [&](auto pr, auto lc, auto sc) {
        hipError_t e = hipSuccess;
        for (int k = 0; k < steps && e == hipSuccess; k++)
            e = batch_step_t<decltype(pr), decltype(lc)::value, decltype(sc)::value>(plan, s);
        return e;
    });
}

hipError_t launch_batch_fused(int prec, int layout, int streaming, const BatchPlan& plan, hipStream_t s) {
    return with_batch_prec(prec, layout, streaming, [&](auto pr, auto lc, auto sc) {
        return batch_fused_t<decltype(pr), decltype(lc)::value, decltype(sc)::value>(plan, s);
    });
}

}  // namespace tdlb
