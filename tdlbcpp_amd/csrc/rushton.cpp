// Host generator of the Rushton-turbine IB point lists (see rushton.h for the
// geometry definition).  Replaces the RushtonTurbinePolarCPP calls of
// tdlbcpp/src/main.cpp:274-303 and :409-426.
#include "rushton.h"

#include <algorithm>
#include <cmath>

namespace tdlb {

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

tdlb_ib_point make_point(double x, double z, int64_t j, double u, double w) {
    tdlb_ib_point p;
    const double fi = std::floor(x + 0.5), fk = std::floor(z + 0.5);
    p.i = (int64_t)fi;
    p.j = j;
    p.k = (int64_t)fk;
    p.i_frac = x - fi;
    p.k_frac = z - fk;
    p.u_delta = u;
    p.v_delta = 0.0;
    p.w_delta = w;
    return p;
}

// ring of points on a circle of radius r around the axis, spacing about 1
void ring(const tdlb_rushton_params& p, double r, int64_t j, double omega, int min_points,
          std::vector<tdlb_ib_point>& out) {
    if (r <= 0.0) {
        out.push_back(make_point(p.center_x, p.center_z, j, 0.0, 0.0));
        return;
    }
    const int n = std::max(min_points, (int)std::ceil(kTwoPi * r));
    for (int a = 0; a < n; a++) {
        const double th = kTwoPi * (double)a / (double)n;
        const double c = std::cos(th), s = std::sin(th);
        out.push_back(make_point(p.center_x + r * c, p.center_z + r * s, j, -omega * (r * s), omega * (r * c)));
    }
}

// stable order by stencil-centre colour (i mod 3, k mod 3): points of one
// colour on one plane overlap only when they share a centre cell
void colour_sort(std::vector<tdlb_ib_point>& v) {
    auto colour = [](const tdlb_ib_point& q) {
        auto m3 = [](int64_t a) { return (int)(((a % 3) + 3) % 3); };
        return m3(q.i) * 3 + m3(q.k);
    };
    std::stable_sort(v.begin(), v.end(),
                     [&](const tdlb_ib_point& a, const tdlb_ib_point& b) { return colour(a) < colour(b); });
}

}  // namespace

void rushton_defaults(int64_t T, int64_t H, tdlb_rushton_params* p) {
    const double Td = (double)T;
    const double D = Td / 3.0;
    p->tank_diameter = T;
    p->height = H;
    p->center_x = (Td - 1.0) / 2.0;
    p->center_z = (Td - 1.0) / 2.0;
    p->tank_radius = (Td - 1.0) / 2.0;
    p->n_baffles = 4;
    p->baffle_width = Td / 10.0;
    p->baffle_angle0 = kTwoPi / 8.0;
    p->disk_plane = std::min<int64_t>(std::max<int64_t>((int64_t)(Td / 3.0), 0), H - 1);
    p->disk_radius = 3.0 * D / 8.0;
    p->hub_radius = D / 8.0;
    p->shaft_radius = D / 12.0;
    p->n_blades = 6;
    p->blade_outer_radius = (double)(int64_t)(Td / 6.0);
    p->blade_inner_radius = p->blade_outer_radius - D / 4.0;
    const int64_t hb = std::max<int64_t>(1, (int64_t)std::llround(D / 5.0));
    p->blade_plane_lo = std::max<int64_t>(0, p->disk_plane - hb / 2);
    p->blade_planes = std::min<int64_t>(hb + 1, H - p->blade_plane_lo);
    p->tip_speed = 0.1;
    p->startup_steps = 30;
}

BladeEnum Rushton::blade_enum() const {
    BladeEnum e;
    e.n_blades = p.n_blades;
    e.n_planes = (int32_t)p.blade_planes;
    e.n_radial = (int32_t)std::floor(p.blade_outer_radius - p.blade_inner_radius + 1e-9) + 1;
    e.n_bcol = blade_colours(p.n_blades, p.blade_inner_radius);
    return e;
}

int64_t Rushton::n_blade_points() const {
    const BladeEnum e = blade_enum();
    return (int64_t)e.n_blades * e.n_planes * e.n_radial;
}

void Rushton::blade_trig(double angle, double* cb, double* sb) const {
    for (int b = 0; b < p.n_blades; b++) {
        const double th = angle + kTwoPi * (double)b / (double)p.n_blades;
        cb[b] = std::cos(th);
        sb[b] = std::sin(th);
    }
}

void Rushton::fixed(std::vector<tdlb_ib_point>& out) const {
    std::vector<tdlb_ib_point> v;
    for (int64_t j = 0; j < p.height; j++) ring(p, p.tank_radius, j, 0.0, 8, v);
    for (int b = 0; b < p.n_baffles; b++) {
        const double th = p.baffle_angle0 + kTwoPi * (double)b / (double)p.n_baffles;
        const double c = std::cos(th), s = std::sin(th);
        for (int64_t j = 0; j < p.height; j++)
            for (double r = p.tank_radius - p.baffle_width; r < p.tank_radius - 0.5; r += 1.0)
                v.push_back(make_point(p.center_x + r * c, p.center_z + r * s, j, 0.0, 0.0));
    }
    colour_sort(v);
    out.insert(out.end(), v.begin(), v.end());
}

void Rushton::rotating_non_updating(double omega, std::vector<tdlb_ib_point>& out) const {
    std::vector<tdlb_ib_point> v;
    // disk: filled, concentric rings on the disk plane
    for (int r = 0; r <= (int)std::floor(p.disk_radius); r++) ring(p, (double)r, p.disk_plane, omega, 1, v);
    // hub over the blade height, shaft from above the disk to the top
    for (int64_t j = p.blade_plane_lo; j < p.blade_plane_lo + p.blade_planes; j++)
        if (j != p.disk_plane) ring(p, p.hub_radius, j, omega, 4, v);
    for (int64_t j = p.blade_plane_lo + p.blade_planes; j < p.height; j++) ring(p, p.shaft_radius, j, omega, 4, v);
    colour_sort(v);
    out.insert(out.end(), v.begin(), v.end());
}

void Rushton::rotating(double angle, double omega, std::vector<tdlb_ib_point>& out) const {
    const BladeEnum e = blade_enum();
    double cb[64], sb[64];
    blade_trig(angle, cb, sb);
    const int64_t n = n_blade_points();
    const size_t base = out.size();
    out.resize(base + n);
    for (int64_t t = 0; t < n; t++) {
        int b, pl, r, cls;
        blade_decode(e, t, b, pl, r, cls);
        blade_point(p.center_x, p.center_z, p.blade_inner_radius, (int32_t)(p.blade_plane_lo + pl), r, cb[b], sb[b],
                    omega, &out[base + t]);
    }
}

void Rushton::solid_mask(int64_t xg, int64_t yg, int64_t zg, int64_t ghost, uint8_t* out) const {
    for (int64_t i = 0; i < xg; i++)
        for (int64_t j = 0; j < yg; j++)
            for (int64_t k = 0; k < zg; k++) {
                const double dx = (double)(i - ghost) - p.center_x, dz = (double)(k - ghost) - p.center_z;
                out[(i * yg + j) * zg + k] = (dx * dx + dz * dz > p.tank_radius * p.tank_radius) ? 1 : 0;
            }
}

double Rushton::step_increment(uint64_t step) const {
    const double full = p.tip_speed / p.blade_outer_radius;
    if (p.startup_steps > 0 && step < (uint64_t)p.startup_steps)
        return 0.5 * full * (1.0 - std::cos(3.14159265358979323846 * (double)step / (double)p.startup_steps));
    return full;
}

}  // namespace tdlb
