// Rushton-turbine stirred tank: immersed-boundary point generator.
//
// The reference takes these lists from tdLBGeometryRushtonTurbineLib
// (RushtonTurbinePolarCPP: generateFixedGeometry, generateRotatingNonUpdatingGeometry,
// generateRotatingGeometry, updateRotatingGeometry, calcThisStepImpellerIncrement,
// getExternalPoints; call sites tdlbcpp/src/main.cpp:274-303,409-426).  That
// library is an empty git submodule in the reference tree, so this generator is
// this project's own definition of the same geometry:
//
//   tank      cylinder of diameter T = grid x, axis along y (vertical) through
//             ((T-1)/2, (T-1)/2) in (i,k); wall points on every j plane
//   baffles   4 flat baffles of width T/10 at 45 + 90n degrees, full height
//   impeller  diameter D = T/3, clearance T/3: disk plane j = T/3
//     disk    filled disk of radius 3D/8 on the disk plane
//     hub     cylinder of radius D/8 over the blade height
//     shaft   cylinder of radius D/12 from the disk plane to the top
//     blades  6 flat radial blades, outer radius int(T/6) (the integer
//             outerRadius the reference feeds to FlowParams::calcNuAndRe_m),
//             length D/4, height D/5 centred on the disk plane
//
// Points are spaced about one lattice unit along every curve.  A point at real
// position (x, z) on integer plane j stores i = floor(x + 0.5), k = floor(z + 0.5)
// and the fractions x - i, z - k in [-1/2, 1/2), so that the 3-point smoothed
// delta of the forcing covers the whole kernel support.
//
// Groups and their order follow main.cpp: fixed (tank, baffles), rotating
// non-updating (disk, hub, shaft: positions fixed, tangential velocity of the
// angular increment they were generated with), rotating (blades, regenerated
// every step).  Static groups are emitted in colour order (i mod 3, k mod 3) of
// their stencil centre so that few points overlap within a run; blades are
// emitted colour-major (blade parity, radial index mod 5), which makes points of
// one colour pairwise non-overlapping whenever the blades of one colour are
// at least 3*sqrt(2) cells apart at the inner radius (checked at bind time; true
// for T = 30 and every T >= 36).
#pragma once

#include <cstdint>
#include <cmath>
#include <vector>

#include "lattice.h"
#include "../../include/tdlb.h"

namespace tdlb {

struct BladeEnum {
    int32_t n_blades, n_planes, n_radial;
    int32_t n_bcol;  // blade colours (blade_colours)
};

// Blade colours: blades b, b + n_bcol, ... share a colour class, so their
// stencils (3 x 3 cells in (i, k)) must not overlap: 2 colours for an even
// blade count, 3 for an odd one, when blades of one colour are more than
// 3 sqrt(2) cells apart at the inner radius; otherwise (small tanks) every
// blade is a colour of its own.
TDLB_HD inline int32_t blade_colours(int32_t n_blades, double r_inner) {
    const int32_t c = (n_blades % 2 == 0) ? 2 : 3;
    if (c >= n_blades) return n_blades > 0 ? n_blades : 1;
    const double sep = 2.0 * r_inner * sin(3.14159265358979323846 * c / n_blades);
    return sep < 3.0 * 1.4142135623730951 + 1e-9 ? n_blades : c;
}

// Number of blade points of colour class c (c = bcol*5 + rmod).
TDLB_HD inline int64_t blade_class_size(const BladeEnum& e, int c) {
    int bcol = c / 5, rmod = c % 5;
    int nb = 0;
    for (int b = bcol; b < e.n_blades; b += e.n_bcol) nb++;
    int nr = (e.n_radial > rmod) ? (e.n_radial - rmod + 4) / 5 : 0;
    return (int64_t)nb * e.n_planes * nr;
}

// Enumeration index -> (blade, plane, radial index), colour-major order:
// colour (blade colour, radial index mod 5), then blade, plane, radial index.
TDLB_HD inline void blade_decode(const BladeEnum& e, int64_t idx, int& b, int& p, int& r, int& cls) {
    int c = 0;
    int64_t sz = blade_class_size(e, 0);
    while (idx >= sz) {
        idx -= sz;
        c++;
        sz = blade_class_size(e, c);
    }
    cls = c;
    int bcol = c / 5, rmod = c % 5;
    int nr = (e.n_radial - rmod + 4) / 5;
    int64_t per_blade = (int64_t)e.n_planes * nr;
    int bi = (int)(idx / per_blade);
    int64_t rem = idx % per_blade;
    b = bcol + bi * e.n_bcol;
    p = (int)(rem / nr);
    r = rmod + 5 * (int)(rem % nr);
}

// One blade point: blade b at orientation (cb, sb), plane j (pre-ghost), radius
// r_inner + r.  Bit-identical on host and device: only +, -, *, floor in double.
TDLB_HD inline void blade_point(double cx, double cz, double r_inner, int32_t j, int r, double cb, double sb,
                                double omega, tdlb_ib_point* out) {
    const double rad = r_inner + (double)r;
    const double x = cx + rad * cb;
    const double z = cz + rad * sb;
    const double fi = floor(x + 0.5);
    const double fk = floor(z + 0.5);
    out->i = (int64_t)fi;
    out->j = j;
    out->k = (int64_t)fk;
    out->i_frac = x - fi;
    out->k_frac = z - fk;
    out->u_delta = -omega * (rad * sb);
    out->v_delta = 0.0;
    out->w_delta = omega * (rad * cb);
}

struct Rushton {
    tdlb_rushton_params p;

    explicit Rushton(const tdlb_rushton_params& prm) : p(prm) {}

    BladeEnum blade_enum() const;
    int64_t n_blade_points() const;
    // blade orientation for a rotation angle
    void blade_trig(double angle, double* cb, double* sb) const;

    void fixed(std::vector<tdlb_ib_point>& out) const;
    void rotating_non_updating(double delta_angle, std::vector<tdlb_ib_point>& out) const;
    void rotating(double angle, double delta_angle, std::vector<tdlb_ib_point>& out) const;
    // cells outside the tank, 1 byte per cell including ghosts (getExternalPoints)
    void solid_mask(int64_t xg, int64_t yg, int64_t zg, int64_t ghost, uint8_t* out) const;
    // angular increment of a step with the start-up ramp (calcThisStepImpellerIncrement)
    double step_increment(uint64_t step) const;
};

void rushton_defaults(int64_t tank_diameter, int64_t height, tdlb_rushton_params* p);

}  // namespace tdlb
