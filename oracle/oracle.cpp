// oracle.cpp -- FP64 CPU restatement of the OpenCafeMol MD step.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
// compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
// bench.py's cpu_baseline leg, and by nothing else.  The product
// (libocm.so) never links or calls it.
//
// What it restates (file:line of the reference, OpenMM semantics where the
// reference hands a Lepton expression to OpenMM 8.0):
//   bonded_*      HarmonicBondForceFieldGenerator.cpp:3-14, HarmonicAngle...cpp:3-15,
//                 GaussianBond...cpp:5-20, GoContact...cpp:5-17,
//                 GaussianDihedral...cpp:5-26, FlexibleLocalAngle...cpp:3-45,
//                 FlexibleLocalDihedral...cpp:28-63
//   pair_*        ExcludedVolume...cpp:23-92, DebyeHuckel...cpp:54-107,
//                 LennardJonesRepulsive...cpp:34-112, LennardJonesAttractive...cpp:38-121
//                 (CustomNonbondedForce: energy 0 at r >= cutoff, exclusions,
//                  interaction groups from InteractionGroup.hpp:17-101)
//   fla_min_energy   FlexibleLocalAngleForceFieldGenerator.cpp:11-19,47-68
//   fld_min_energy   FlexibleLocalDihedralForceFieldGenerator.cpp:7-26
//   natural_spline   OpenMM Continuous1DFunction (natural cubic spline fit)
//   eval_cells       oracle_eval through cell lists (every bead or a sampled
//                    subset; large systems), pinned against oracle_eval
//   baoab            the Langevin step this library implements (BAOAB splitting
//                    with Philox4x32-10 noise); the reference's own integrator is
//                    OpenMM's LangevinIntegrator (IntegratorGenerator.cpp:5-14),
//                    whose RNG stream cannot be reproduced without OpenMM.
//
// How each function is pinned (tests/test_oracle.py): closed-form single terms,
// central finite differences of the energy, Newton's third law, the sh3 sample's
// native geometry (its v0 values are the native distances and dihedrals),
// Random123's published Philox known-answer vectors, numpy's natural spline
// (scipy.interpolate.CubicSpline bc_type='natural'), brute-force minima of the
// spline/Fourier sums, and BAOAB energy conservation at zero friction.

#include "../include/ocm.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstddef>
#include <cstring>
#include <set>
#include <utility>
#include <vector>

namespace {

constexpr double kPi = 3.1415926535897932385;

struct V3 { double x, y, z; };
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

struct Box {
    bool periodic;
    double L[3], invL[3];
};

// Minimum-image displacement x[j]-x[i]; same operation order as the device
// code (one rounded multiply, rint, one fused multiply-add per component).
inline V3 disp(const double* x, int i, int j, const Box& b, bool use_pbc) {
    double d[3];
    for (int c = 0; c < 3; ++c) {
        d[c] = x[3 * j + c] - x[3 * i + c];
        if (use_pbc && b.periodic) d[c] = std::fma(-b.L[c], std::rint(d[c] * b.invL[c]), d[c]);
    }
    return {d[0], d[1], d[2]};
}
inline double r2_of(V3 d) { return std::fma(d.z, d.z, std::fma(d.y, d.y, d.x * d.x)); }

// Natural cubic spline second derivatives for uniformly spaced knots
// (tridiagonal solve, Thomas algorithm), the fit OpenMM's Continuous1DFunction
// applies to the 10 FLA y-values.
void natural_spline_impl(const double* y, int n, double h, double* y2) {
    std::vector<double> c(n, 0.0), d(n, 0.0);
    y2[0] = 0.0;
    y2[n - 1] = 0.0;
    if (n < 3) { for (int i = 0; i < n; ++i) y2[i] = 0.0; return; }
    // rows i=1..n-2: h*y2[i-1] + 4h*y2[i] + h*y2[i+1] = 6 (y[i+1]-2y[i]+y[i-1]) / h
    std::vector<double> diag(n), rhs(n);
    for (int i = 1; i < n - 1; ++i) {
        diag[i] = 4.0 * h;
        rhs[i] = 6.0 * (y[i + 1] - 2.0 * y[i] + y[i - 1]) / h;
    }
    // forward sweep
    for (int i = 2; i < n - 1; ++i) {
        const double m = h / diag[i - 1];
        diag[i] -= m * h;
        rhs[i] -= m * rhs[i - 1];
    }
    y2[n - 2] = rhs[n - 2] / diag[n - 2];
    for (int i = n - 3; i >= 1; --i) y2[i] = (rhs[i] - h * y2[i + 1]) / diag[i];
}

// Value and derivative of the natural spline at t (caller ensures the range).
void spline_eval(const double* y, const double* y2, int n, double tmin, double h,
                 double t, double* val, double* der) {
    int lo = static_cast<int>(std::floor((t - tmin) / h));
    lo = std::max(0, std::min(n - 2, lo));
    const double xl = tmin + lo * h, xu = tmin + (lo + 1) * h;
    const double a = (xu - t) / h, b = (t - xl) / h;
    *val = a * y[lo] + b * y[lo + 1] +
           ((a * a * a - a) * y2[lo] + (b * b * b - b) * y2[lo + 1]) * h * h / 6.0;
    *der = (y[lo + 1] - y[lo]) / h +
           ((1.0 - 3.0 * a * a) * y2[lo] + (3.0 * b * b - 1.0) * y2[lo + 1]) * h / 6.0;
}

inline void addf(double* F, int i, V3 f) {
    if (!F) return;
    F[3 * i + 0] += f.x; F[3 * i + 1] += f.y; F[3 * i + 2] += f.z;
}

// ---- angle theta at j and its gradient (atan2 form, cross-product gradient)
double angle_grad(const double* x, int i, int j, int k, const Box& b, bool pbc,
                  V3* gi, V3* gj, V3* gk) {
    const V3 a = disp(x, j, i, b, pbc);   // x_i - x_j
    const V3 c = disp(x, j, k, b, pbc);   // x_k - x_j
    const V3 p = cross(a, c);
    const double pn = norm(p);
    const double th = std::atan2(pn, dot(a, c));
    if (pn < 1e-300) { *gi = *gj = *gk = {0, 0, 0}; return th; }
    *gi = mul(cross(a, p), 1.0 / (dot(a, a) * pn));
    *gk = mul(cross(p, c), 1.0 / (dot(c, c) * pn));
    *gj = mul(add(*gi, *gk), -1.0);
    return th;
}

// ---- IUPAC dihedral phi = atan2(|b2| b1.(b2 x b3), (b1 x b2).(b2 x b3))
double dihedral_grad(const double* x, int i, int j, int k, int l, const Box& b, bool pbc,
                     V3* gi, V3* gj, V3* gk, V3* gl) {
    const V3 b1 = disp(x, i, j, b, pbc);
    const V3 b2 = disp(x, j, k, b, pbc);
    const V3 b3 = disp(x, k, l, b, pbc);
    const V3 n1 = cross(b1, b2), n2 = cross(b2, b3);
    const double b2n = norm(b2);
    const double phi = std::atan2(b2n * dot(b1, n2), dot(n1, n2));
    const double n1s = dot(n1, n1), n2s = dot(n2, n2);
    if (n1s < 1e-300 || n2s < 1e-300 || b2n < 1e-300) {
        *gi = *gj = *gk = *gl = {0, 0, 0};
        return phi;
    }
    *gi = mul(n1, -b2n / n1s);
    *gl = mul(n2, b2n / n2s);
    const double f1 = dot(b1, b2) / (b2n * b2n), f3 = dot(b3, b2) / (b2n * b2n);
    // middle beads (Blondel & Karplus, J. Comput. Chem. 17, 1132 (1996)):
    // d phi/d x_j = -(1 + f1) d phi/d x_i + f3 d phi/d x_l, and the four sum to 0
    *gj = sub(mul(*gl, f3), mul(*gi, 1.0 + f1));
    *gk = sub(mul(*gi, f1), mul(*gl, 1.0 + f3));
    return phi;
}

Box make_box(const ocm_system_desc* s) {
    Box b{};
    b.periodic = s->use_periodic_box != 0;
    for (int c = 0; c < 3; ++c) {
        b.L[c] = s->box[c];
        b.invL[c] = b.periodic ? 1.0 / s->box[c] : 0.0;
    }
    return b;
}

// ------------------------------------------------------------- bonded terms
double eval_bonded(const ocm_force_desc& f, const double* x, const Box& box, double* F) {
    const bool pbc = f.use_periodic != 0;
    double E = 0.0;
    for (int t = 0; t < f.n_terms; ++t) {
        switch (f.kind) {
        case OCM_HARMONIC_BOND:
        case OCM_GAUSSIAN_BOND:
        case OCM_GO_CONTACT:
        case OCM_CAPPED_GO_CONTACT: {
            const int i = f.indices[2 * t], j = f.indices[2 * t + 1];
            const V3 d = disp(x, i, j, box, pbc);
            const double r = norm(d);
            double e = 0, dedr = 0;
            if (f.kind == OCM_HARMONIC_BOND) {  // params {r0, k}
                const double r0 = f.params[2 * t], k = f.params[2 * t + 1];
                e = 0.5 * k * (r - r0) * (r - r0);
                dedr = k * (r - r0);
            } else if (f.kind == OCM_GAUSSIAN_BOND) {  // {k, v0, sigma}
                const double k = f.params[3 * t], v0 = f.params[3 * t + 1], s = f.params[3 * t + 2];
                e = k * std::exp(-(r - v0) * (r - v0) / (2.0 * s * s));
                dedr = -e * (r - v0) / (s * s);
            } else if (f.kind == OCM_GO_CONTACT) {  // {k, r0}
                const double k = f.params[2 * t], r0 = f.params[2 * t + 1];
                const double q2 = (r0 / r) * (r0 / r);
                const double q10 = q2 * q2 * q2 * q2 * q2, q12 = q10 * q2;
                e = k * (5.0 * q12 - 6.0 * q10);
                dedr = 60.0 * k * (q10 - q12) / r;
            } else {
                // CappedGoContactForceFieldGenerator.cpp:5-26, params {k, r0, cap}:
                // cap_val * cap + (1 - cap_val) * lj, cap_val = step(cap - r / r0),
                // cap = k slope (r / r0) + k intercept
                const double k = f.params[3 * t], r0 = f.params[3 * t + 1], c = f.params[3 * t + 2];
                const double slope = -60.0 * (std::pow(1.0 / c, 13.0) - std::pow(1.0 / c, 11.0));
                const double intercept = 65.0 * std::pow(1.0 / c, 12.0) - 66.0 * std::pow(1.0 / c, 10.0);
                if (c - r / r0 >= 0.0) {
                    e = k * slope * (r / r0) + k * intercept;
                    dedr = k * slope / r0;
                } else {
                    const double q = r0 / r;
                    e = k * (5.0 * std::pow(q, 12.0) - 6.0 * std::pow(q, 10.0));
                    dedr = k * (-60.0 * std::pow(q, 12.0) + 60.0 * std::pow(q, 10.0)) / r;
                }
            }
            E += e;
            const V3 fi = mul(d, dedr / r);  // F_i = dE/dr * d/r with d = x_j - x_i
            addf(F, i, fi);
            addf(F, j, mul(fi, -1.0));
            break;
        }
        case OCM_HARMONIC_ANGLE:
        case OCM_FLEXIBLE_LOCAL_ANGLE: {
            const int i = f.indices[3 * t], j = f.indices[3 * t + 1], k = f.indices[3 * t + 2];
            V3 gi, gj, gk;
            const double th = angle_grad(x, i, j, k, box, pbc, &gi, &gj, &gk);
            double e = 0, dedt = 0;
            if (f.kind == OCM_HARMONIC_ANGLE) {  // {theta0, k}
                const double t0 = f.params[2 * t], kk = f.params[2 * t + 1];
                e = 0.5 * kk * (th - t0) * (th - t0);
                dedt = kk * (th - t0);
            } else {  // params {k}; globals y[10], tmin, tmax, Emin
                const double kk = f.params[t];
                const double* y = f.globals;
                const double tmin = f.globals[10], tmax = f.globals[11], emin = f.globals[12];
                const double h = (tmax - tmin) / 9.0;
                double y2[10];
                natural_spline_impl(y, 10, h, y2);
                double sv = 0, sd = 0;
                if (th >= tmin && th <= tmax) spline_eval(y, y2, 10, tmin, h, th, &sv, &sd);
                double u = sv - emin, du = sd;
                if (tmin - th >= 0.0) { u += 30.0 * (tmin - th) + y[0]; du -= 30.0; }
                if (th - tmax >= 0.0) { u += 30.0 * (th - tmax) + y[9]; du += 30.0; }
                e = kk * u;
                dedt = kk * du;
            }
            E += e;
            addf(F, i, mul(gi, -dedt));
            addf(F, j, mul(gj, -dedt));
            addf(F, k, mul(gk, -dedt));
            break;
        }
        case OCM_GAUSSIAN_DIHEDRAL:
        case OCM_FLEXIBLE_LOCAL_DIHEDRAL:
        case OCM_COSINE_DIHEDRAL: {
            const int i = f.indices[4 * t], j = f.indices[4 * t + 1];
            const int k = f.indices[4 * t + 2], l = f.indices[4 * t + 3];
            V3 gi, gj, gk, gl;
            const double phi = dihedral_grad(x, i, j, k, l, box, pbc, &gi, &gj, &gk, &gl);
            double e = 0, dedp = 0;
            if (f.kind == OCM_GAUSSIAN_DIHEDRAL) {  // {k, phi0, sigma}
                const double kk = f.params[3 * t], p0 = f.params[3 * t + 1], s = f.params[3 * t + 2];
                const double dt = phi - p0;
                const double dtp = dt - std::floor((dt + kPi) / (2.0 * kPi)) * (2.0 * kPi);
                e = kk * std::exp(-dtp * dtp / (2.0 * s * s));
                dedp = -e * dtp / (s * s);
            } else if (f.kind == OCM_COSINE_DIHEDRAL) {
                // CosineDihedralForceFieldGenerator.cpp:7, params {k, t0, n}:
                // k (1 - cos(n (theta - t0)))
                const double kk = f.params[3 * t], t0 = f.params[3 * t + 1], nn = f.params[3 * t + 2];
                e = kk * (1.0 - std::cos(nn * (phi - t0)));
                dedp = kk * nn * std::sin(nn * (phi - t0));
            } else {  // {k}; globals c,a1,b1,a2,b2,a3,b3,Emin
                const double kk = f.params[t];
                const double* g = f.globals;
                double u = g[0] - g[7], du = 0.0;
                for (int n = 1; n <= 3; ++n) {
                    const double a = g[2 * n - 1], bb = g[2 * n];
                    u += a * std::cos(n * phi) + bb * std::sin(n * phi);
                    du += n * (-a * std::sin(n * phi) + bb * std::cos(n * phi));
                }
                e = kk * u;
                dedp = kk * du;
            }
            E += e;
            addf(F, i, mul(gi, -dedp));
            addf(F, j, mul(gj, -dedp));
            addf(F, k, mul(gk, -dedp));
            addf(F, l, mul(gl, -dedp));
            break;
        }
        default:
            break;
        }
    }
    return E;
}

// ---------------------------------------------------------- nonbonded terms
struct PairCtx {
    const ocm_force_desc* f;
    int n;
    std::set<std::pair<int, int>> excl;
};

inline bool participates(const ocm_force_desc& f, int i) {
    return f.category == nullptr || f.category[i] >= 0;
}
inline bool allowed(const ocm_force_desc& f, int i, int j) {
    if (f.category == nullptr || f.category_allow == nullptr) return true;
    return f.category_allow[f.category[i] * f.n_categories + f.category[j]] != 0;
}

PairCtx make_pair_ctx(const ocm_force_desc& f, int n) {
    PairCtx c{&f, n, {}};
    for (int64_t e = 0; e < f.n_exclusions; ++e) {
        int a = f.exclusions[2 * e], b = f.exclusions[2 * e + 1];
        if (a > b) std::swap(a, b);
        c.excl.insert({a, b});
    }
    return c;
}

// Energy and dE/dr of one pair of the potential `f` at distance r.
void pair_energy(const ocm_force_desc& f, int i, int j, double r, double* e, double* dedr) {
    const double* pp = f.particle_params;
    switch (f.kind) {
    case OCM_EXCLUDED_VOLUME: {
        const double eps = f.globals[0];
        const double s = pp[i] + pp[j];
        const double s12 = std::pow(s, 12);
        const double corr = std::pow(1.0 / f.cutoff, 12);
        *e = eps * s12 * (std::pow(1.0 / r, 12) - corr);
        *dedr = -12.0 * eps * s12 * std::pow(1.0 / r, 13);
        break;
    }
    case OCM_DEBYE_HUCKEL: {
        const double K = f.globals[0], lam = f.globals[1], corr = f.globals[2];
        const double qq = pp[i] * pp[j];
        const double ex = std::exp(-r / lam);
        *e = qq * K * (ex / r - corr);
        *dedr = -qq * K * ex * (1.0 / (r * r) + 1.0 / (lam * r));
        break;
    }
    case OCM_LJ_REPULSIVE:
    case OCM_LJ_ATTRACTIVE: {
        const double eb = 0.5 * (pp[2 * i] + pp[2 * j]);
        const double sb = 0.5 * (pp[2 * i + 1] + pp[2 * j + 1]);
        const double th = sb * std::pow(2.0, 1.0 / 6.0);
        const double sr = sb / r, s6 = std::pow(sr, 6), s12 = s6 * s6;
        double u = 0, du = 0;
        if (f.kind == OCM_LJ_REPULSIVE) {
            if (th - r >= 0.0) {
                u = eb * (4.0 * (s12 - s6) + 1.0);
                du = eb * 4.0 * (-12.0 * s12 + 6.0 * s6) / r;
            }
        } else {
            if (r - th >= 0.0) {
                u += eb * 4.0 * (s12 - s6);
                du += eb * 4.0 * (-12.0 * s12 + 6.0 * s6) / r;
            }
            if (th - r >= 0.0) u -= eb;
        }
        const double rsw = f.switch_distance, rc = f.cutoff;
        if (rsw > 0.0 && r > rsw) {
            const double xx = (r - rsw) / (rc - rsw);
            const double S = 1.0 - 10.0 * xx * xx * xx + 15.0 * xx * xx * xx * xx - 6.0 * std::pow(xx, 5);
            const double dS = (-30.0 * xx * xx + 60.0 * xx * xx * xx - 30.0 * xx * xx * xx * xx) / (rc - rsw);
            du = du * S + u * dS;
            u = u * S;
        }
        *e = u;
        *dedr = du;
        break;
    }
    case OCM_WCA: {
        // WeeksChandlerAndersenForceFieldGenerator.cpp:34-40, npp {sigma, eps}
        const double sb = (pp[2 * i] + pp[2 * j]) * 0.5;
        const double eb = std::sqrt(pp[2 * i + 1] * pp[2 * j + 1]);
        const double sr = sb / r, s6 = std::pow(sr, 6), s12 = s6 * s6;
        if (sb * std::pow(2.0, 1.0 / 6.0) - r >= 0.0) {
            *e = 4.0 * eb * (s12 - s6) + eb;
            *dedr = eb * 4.0 * (-12.0 * s12 + 6.0 * s6) / r;
        } else {
            *e = 0; *dedr = 0;
        }
        break;
    }
    case OCM_UNIFORM_WCA: {
        // UniformWeeksChandlerAndersenForceFieldGenerator.cpp:162-167, globals {sigma, eps}
        const double sg = f.globals[0], ep = f.globals[1];
        const double sr = sg / r, s6 = std::pow(sr, 6), s12 = s6 * s6;
        *e = 4.0 * ep * (s12 - s6 + 0.25);
        *dedr = ep * 4.0 * (-12.0 * s12 + 6.0 * s6) / r;
        break;
    }
    case OCM_UNIFORM_LJ_ATTRACTIVE: {
        // UniformLennardJonesAttractiveForceFieldGenerator.cpp:145-151 and its
        // switch from 0.9 cutoff (:186-187), globals {sigma, eps}
        const double sg = f.globals[0], ep = f.globals[1];
        const double th = sg * std::pow(2.0, 1.0 / 6.0);
        const double sr = sg / r, s6 = std::pow(sr, 6), s12 = s6 * s6;
        double u = 0, du = 0;
        if (r - th >= 0.0) {
            u += ep * 4.0 * (s12 - s6);
            du += ep * 4.0 * (-12.0 * s12 + 6.0 * s6) / r;
        }
        if (th - r >= 0.0) u -= ep;
        const double rsw = f.switch_distance, rc = f.cutoff;
        if (rsw > 0.0 && r > rsw) {
            const double xx = (r - rsw) / (rc - rsw);
            const double S = 1.0 - 10.0 * xx * xx * xx + 15.0 * xx * xx * xx * xx - 6.0 * std::pow(xx, 5);
            const double dS = (-30.0 * xx * xx + 60.0 * xx * xx * xx - 30.0 * xx * xx * xx * xx) / (rc - rsw);
            du = du * S + u * dS;
            u = u * S;
        }
        *e = u;
        *dedr = du;
        break;
    }
    case OCM_ISOLF_ATTRACTIVE: {
        // iSoLFAttractiveForceFieldGenerator.cpp:43-51, npp {sigma, eps, omega}
        const double sb = (pp[3 * i] + pp[3 * j]) * 0.5;
        const double eb = std::sqrt(pp[3 * i + 1] * pp[3 * j + 1]);
        const double om = (pp[3 * i + 2] + pp[3 * j + 2]) * 0.5;
        const double base = sb * std::pow(2.0, 1.0 / 6.0);
        const double step_in = base - r >= 0.0 ? 1.0 : 0.0;
        const double step_tail = (r - base >= 0.0 ? 1.0 : 0.0) * (base + om - r >= 0.0 ? 1.0 : 0.0);
        const double c = std::cos(0.5 * kPi * (r - base) / om);
        *e = -eb * (step_in + step_tail * c * c);
        *dedr = eb * step_tail * 2.0 * c * std::sin(0.5 * kPi * (r - base) / om) * 0.5 * kPi / om;
        break;
    }
    default:
        *e = 0; *dedr = 0;
    }
}

double eval_pairs(const ocm_force_desc& f, int n, const double* x, const Box& box, double* F) {
    const PairCtx c = make_pair_ctx(f, n);
    const bool pbc = f.use_periodic != 0;
    const double rc2 = f.cutoff * f.cutoff;
    double E = 0.0;
    for (int i = 0; i < n; ++i) {
        if (!participates(f, i)) continue;
        for (int j = i + 1; j < n; ++j) {
            if (!participates(f, j) || !allowed(f, i, j)) continue;
            if (c.excl.count({i, j})) continue;
            const V3 d = disp(x, i, j, box, pbc);
            const double r2 = r2_of(d);
            if (!(r2 < rc2)) continue;
            const double r = std::sqrt(r2);
            double e, dedr;
            pair_energy(f, i, j, r, &e, &dedr);
            E += e;
            const V3 fi = mul(d, dedr / r);
            addf(F, i, fi);
            addf(F, j, mul(fi, -1.0));
        }
    }
    return E;
}

bool is_bonded(int kind) { return kind >= OCM_HARMONIC_BOND && kind <= OCM_COSINE_DIHEDRAL; }

// ------------------------------------------------------------------ Philox
inline void philox_round(uint32_t* c, const uint32_t* k) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c[0];
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n1 = lo1;
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    const uint32_t n3 = lo0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
}

void philox10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        philox_round(c, k);
        if (r < 9) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
    }
    for (int i = 0; i < 4; ++i) out[i] = c[i];
}

// Three standard normals for (seed, step, bead, stream): Box-Muller on the
// four 32-bit Philox words, each mapped to (w + 0.5) / 2^32.
void gauss3(uint64_t seed, uint64_t step, uint32_t bead, uint32_t stream, double g[3]) {
    const uint32_t ctr[4] = {static_cast<uint32_t>(step), static_cast<uint32_t>(step >> 32), bead, stream};
    const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
    uint32_t w[4];
    philox10(ctr, key, w);
    const double s = 2.3283064365386963e-10;  // 2^-32
    const double u0 = (w[0] + 0.5) * s, u1 = (w[1] + 0.5) * s;
    const double u2 = (w[2] + 0.5) * s, u3 = (w[3] + 0.5) * s;
    const double r0 = std::sqrt(-2.0 * std::log(u0)), r1 = std::sqrt(-2.0 * std::log(u2));
    g[0] = r0 * std::cos(2.0 * kPi * u1);
    g[1] = r0 * std::sin(2.0 * kPi * u1);
    g[2] = r1 * std::cos(2.0 * kPi * u3);
}

}  // namespace

extern "C" {

// Total potential energy per force group (E_by_group[32], kJ/mol) and forces
// F (n*3, kJ/mol/nm; may be NULL) of the groups selected in `groups`.
int oracle_eval(const ocm_system_desc* s, const double* x, uint32_t groups, double* F,
                double* E_by_group) {
    const Box box = make_box(s);
    const int n = s->n_particles;
    if (F) std::memset(F, 0, sizeof(double) * 3 * n);
    if (E_by_group) std::memset(E_by_group, 0, sizeof(double) * 32);
    for (int fi = 0; fi < s->n_forces; ++fi) {
        const ocm_force_desc& f = s->forces[fi];
        if (!((groups >> f.group) & 1u)) continue;
        const double e = is_bonded(f.kind) ? eval_bonded(f, x, box, F) : eval_pairs(f, n, x, box, F);
        if (E_by_group) E_by_group[f.group] += e;
    }
    return 0;
}

// Pairs (i<j) of nonbonded force `force_index` with squared minimum-image
// distance below radius^2 that participate, are allowed and not excluded.
int64_t oracle_neighbor_pairs(const ocm_system_desc* s, int32_t force_index, const double* x,
                              double radius, int32_t* out, int64_t cap) {
    const ocm_force_desc& f = s->forces[force_index];
    const Box box = make_box(s);
    const PairCtx c = make_pair_ctx(f, s->n_particles);
    const double rr = radius * radius;
    int64_t cnt = 0;
    for (int i = 0; i < s->n_particles; ++i) {
        if (!participates(f, i)) continue;
        for (int j = i + 1; j < s->n_particles; ++j) {
            if (!participates(f, j) || !allowed(f, i, j) || c.excl.count({i, j})) continue;
            if (!(r2_of(disp(x, i, j, box, f.use_periodic != 0)) < rr)) continue;
            if (cnt < cap && out) { out[2 * cnt] = i; out[2 * cnt + 1] = j; }
            ++cnt;
        }
    }
    return cnt;
}

void oracle_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    philox10(ctr, key, out);
}

void oracle_gauss3(uint64_t seed, uint64_t step, uint32_t bead, uint32_t stream, double* g) {
    gauss3(seed, step, bead, stream, g);
}

void oracle_natural_spline(const double* y, int32_t n, double h, double* y2) {
    natural_spline_impl(y, n, h, y2);
}

// The reference's own evaluator of its y''-table: interval search over the
// rounded knot list and the cubic-spline formula, swept in 1e-4 rad steps
// from tmin while th < tmax (FlexibleLocalAngleForceFieldGenerator.cpp:11-19,47-68).
double oracle_fla_min_energy(const double* y, const double* y2tab, double tmin, double tmax) {
    static const double knots[10] = {1.30900, 1.48353, 1.65806, 1.83260, 2.00713,
                                     2.18166, 2.35619, 2.53073, 2.70526, 2.87979};
    const double dth = (tmax - tmin) / 9.0, rdth = 1.0 / dth;
    double emin = y[0];
    double th = tmin;
    while (th < tmax) {
        long n = static_cast<long>(std::floor((th - tmin) * rdth));
        n = std::min<long>(8, n);
        if (knots[n + 1] < th) ++n;
        else if (th < knots[n]) --n;
        const double a = (knots[n + 1] - th) * rdth, b = (th - knots[n]) * rdth;
        const double e = a * y[n] + b * y[n + 1] +
                         ((a * a * a - a) * y2tab[n] + (b * b * b - b) * y2tab[n + 1]) * dth * dth / 6.0;
        emin = std::min(emin, e);
        th += 1e-4;
    }
    return emin;
}

// FlexibleLocalDihedralForceFieldGenerator.cpp:7-26: min over phi in
// [-pi, pi) in 1e-4 steps of c + sum a_n cos(n phi) + b_n sin(n phi).
double oracle_fld_min_energy(const double* t7) {
    double phi = -kPi, emin = 1.79769313486231570815e308;
    while (phi < kPi) {
        const double s1 = std::sin(phi), c1 = std::cos(phi);
        const double ss = s1 * s1, cc = c1 * c1;
        const double s2 = 2 * s1 * c1, c2 = cc - ss;
        const double s3 = s1 * (3 - 4 * ss), c3 = c1 * (4 * cc - 3);
        emin = std::min(emin, t7[1] * c1 + t7[2] * s1 + t7[3] * c2 + t7[4] * s2 +
                                  t7[5] * c3 + t7[6] * s3 + t7[0]);
        phi += 1e-4;
    }
    return emin;
}

// BAOAB Langevin steps in FP64 with the Philox noise stream 0, starting at
// step index `step0`.  x, v are n*3 (nm, nm/ps) and are advanced in place.
// The force at the initial x is recomputed first.
int oracle_baoab(const ocm_system_desc* s, const ocm_integrator_desc* it, double* x, double* v,
                 int64_t step0, int64_t nsteps) {
    const int n = s->n_particles;
    const double kB = 1.380649e-23 * 6.02214076e23 / 1000.0;  // kJ/mol/K
    const double kT = kB * it->temperature;
    const double h = 0.5 * it->dt;
    const double c1 = std::exp(-it->friction * it->dt);
    const double c2 = 1.0 - c1 * c1;
    std::vector<double> F(3 * n), im(n);
    for (int i = 0; i < n; ++i) im[i] = s->masses[i] > 0 ? 1.0 / s->masses[i] : 0.0;
    oracle_eval(s, x, 0xFFFFFFFFu, F.data(), nullptr);
    for (int64_t st = 0; st < nsteps; ++st) {
        const uint64_t sidx = static_cast<uint64_t>(step0 + st);
        for (int i = 0; i < n; ++i) {
            if (im[i] == 0.0) continue;
            double g[3];
            gauss3(it->seed, sidx, static_cast<uint32_t>(i), 0u, g);
            const double sig = std::sqrt(kT * c2 * im[i]);
            for (int c = 0; c < 3; ++c) {
                double vv = v[3 * i + c] + h * F[3 * i + c] * im[i];
                double xx = x[3 * i + c] + h * vv;
                vv = c1 * vv + sig * g[c];
                xx = xx + h * vv;
                v[3 * i + c] = vv;
                x[3 * i + c] = xx;
            }
        }
        oracle_eval(s, x, 0xFFFFFFFFu, F.data(), nullptr);
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) v[3 * i + c] += h * F[3 * i + c] * im[i];
    }
    return 0;
}

double oracle_kinetic(const ocm_system_desc* s, const double* v) {
    double ke = 0.0;
    for (int i = 0; i < s->n_particles; ++i)
        ke += 0.5 * s->masses[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    return ke;
}

}  // extern "C"

extern "C" {
// Sizes and field offsets of the ABI structs as this compiler lays them out;
// tests compare them with the ctypes mirror.
int32_t oracle_abi_layout(int64_t* out, int32_t cap) {
    const int64_t v[] = {
        (int64_t)sizeof(ocm_force_desc), (int64_t)offsetof(ocm_force_desc, indices),
        (int64_t)offsetof(ocm_force_desc, globals), (int64_t)offsetof(ocm_force_desc, particle_params),
        (int64_t)offsetof(ocm_force_desc, cutoff), (int64_t)offsetof(ocm_force_desc, n_exclusions),
        (int64_t)offsetof(ocm_force_desc, exclusions), (int64_t)sizeof(ocm_system_desc),
        (int64_t)offsetof(ocm_system_desc, box), (int64_t)offsetof(ocm_system_desc, forces),
        (int64_t)sizeof(ocm_integrator_desc), (int64_t)sizeof(ocm_options),
        (int64_t)offsetof(ocm_options, skin), (int64_t)sizeof(ocm_state),
        (int64_t)offsetof(ocm_state, energy_by_group), (int64_t)offsetof(ocm_state, step),
        (int64_t)sizeof(ocm_server_options), (int64_t)offsetof(ocm_server_options, hbm_budget_bytes),
        (int64_t)offsetof(ocm_server_options, quantum_steps), (int64_t)sizeof(ocm_server_stats),
        (int64_t)offsetof(ocm_server_stats, latency_p50_ms), (int64_t)offsetof(ocm_server_stats, budget_bytes),
        (int64_t)sizeof(ocm_job_info), (int64_t)offsetof(ocm_job_info, steps_done),
        (int64_t)offsetof(ocm_job_info, end_ms)};
    const int32_t n = (int32_t)(sizeof(v) / sizeof(v[0]));
    for (int32_t i = 0; i < n && i < cap; ++i) out[i] = v[i];
    return n;
}
}

// ----------------------------------------------------------------------------
// Cell-list CPU step (CPU baseline): the same BAOAB and the same per-pair
// formulas as above, with nonbonded pairs found through a cell list and forces
// gathered per bead (full neighbor search, OpenMP over beads).  Used only to
// time a CPU implementation of the step beside the GPU; parity is checked with
// oracle_eval / oracle_baoab.
namespace {

struct CellPairs {
    const ocm_force_desc* f;
    std::vector<std::vector<int>> excl;   // sorted partners per bead
};

CellPairs make_cell_pairs(const ocm_force_desc& f, int n) {
    CellPairs cp{&f, std::vector<std::vector<int>>(n)};
    for (int64_t e = 0; e < f.n_exclusions; ++e) {
        const int a = f.exclusions[2 * e], b = f.exclusions[2 * e + 1];
        if (a == b) continue;
        cp.excl[a].push_back(b);
        cp.excl[b].push_back(a);
    }
    for (auto& v2 : cp.excl) {
        std::sort(v2.begin(), v2.end());
        v2.erase(std::unique(v2.begin(), v2.end()), v2.end());
    }
    return cp;
}

// Nonbonded force f through a cell list: the same pairs (participating,
// allowed, not excluded, minimum-image r^2 < cutoff^2 by r2_of) and the same
// per-pair formulas (pair_energy) as eval_pairs, found from cells of edge >=
// cutoff.  Forces are gathered per bead: on every bead (rows == NULL, F n*3
// indexed by bead) or on beads rows[0..m) (F m*3 in that order), added to F.
// Returns the energy of every pair once (i < j) when want_energy (rows must
// then be NULL), else 0.  OpenMP over the rows.
double pair_cells(const ocm_system_desc* s, const CellPairs& cp, const double* x, double* F, const int* rows,
                  int m, int nthreads, bool want_energy) {
    const ocm_force_desc& f = *cp.f;
    const int n = s->n_particles;
    const Box box = make_box(s);
    const bool pbc = f.use_periodic != 0 && box.periodic;
    const double rc = f.cutoff, rc2 = rc * rc;
    double lo[3], ext[3];
    int nc[3];
    for (int c = 0; c < 3; ++c) {
        if (pbc) {
            lo[c] = 0.0;
            ext[c] = box.L[c];
        } else {
            double mn = 1e300, mx = -1e300;
            for (int i = 0; i < n; ++i) { mn = std::min(mn, x[3 * i + c]); mx = std::max(mx, x[3 * i + c]); }
            lo[c] = mn;
            ext[c] = std::max(mx - mn, 1e-9) * (1.0 + 1e-12);
        }
        nc[c] = std::max(1, static_cast<int>(ext[c] / rc));
        if (pbc && nc[c] < 3) nc[c] = 1;
        nc[c] = std::min(nc[c], 1024);
    }
    const long long ncell = static_cast<long long>(nc[0]) * nc[1] * nc[2];
    std::vector<int> cell_of(n, -1), head(static_cast<size_t>(ncell), -1), next(n, -1);
    for (int i = n - 1; i >= 0; --i) {
        if (!participates(f, i)) continue;
        int ci[3];
        for (int c = 0; c < 3; ++c) {
            double u = (x[3 * i + c] - lo[c]) / ext[c];
            if (pbc) u -= std::floor(u);
            int k = static_cast<int>(u * nc[c]);
            ci[c] = std::max(0, std::min(nc[c] - 1, k));
        }
        const int cell = (ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
        cell_of[i] = cell;
        next[i] = head[cell];
        head[cell] = i;
    }
    const int nrows = rows ? m : n;
    double E = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64) reduction(+ : E)
    for (int q = 0; q < nrows; ++q) {
        const int i = rows ? rows[q] : q;
        if (i < 0 || i >= n || cell_of[i] < 0) continue;
        const int cell = cell_of[i];
        const int cx = cell % nc[0], cy = (cell / nc[0]) % nc[1], cz = cell / (nc[0] * nc[1]);
        double acc[3] = {0, 0, 0};
        for (int dz = -1; dz <= 1; ++dz) {
            if (nc[2] == 1 && dz != 0) continue;
            int z = cz + dz;
            if (pbc) z = (z + nc[2]) % nc[2]; else if (z < 0 || z >= nc[2]) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                if (nc[1] == 1 && dy != 0) continue;
                int y = cy + dy;
                if (pbc) y = (y + nc[1]) % nc[1]; else if (y < 0 || y >= nc[1]) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    if (nc[0] == 1 && dx != 0) continue;
                    int xx = cx + dx;
                    if (pbc) xx = (xx + nc[0]) % nc[0]; else if (xx < 0 || xx >= nc[0]) continue;
                    for (int j = head[(z * nc[1] + y) * nc[0] + xx]; j >= 0; j = next[j]) {
                        if (j == i || !allowed(f, i, j)) continue;
                        const auto& ex = cp.excl[i];
                        if (!ex.empty() && std::binary_search(ex.begin(), ex.end(), j)) continue;
                        // displacement from the lower index, as eval_pairs takes it
                        const int a = std::min(i, j), b = std::max(i, j);
                        const V3 d = disp(x, a, b, box, pbc);
                        const double r2 = r2_of(d);
                        if (!(r2 < rc2)) continue;
                        const double r = std::sqrt(r2);
                        double e, dedr;
                        pair_energy(f, a, b, r, &e, &dedr);
                        const double sgn = i == a ? 1.0 : -1.0;
                        acc[0] += sgn * d.x * dedr / r;
                        acc[1] += sgn * d.y * dedr / r;
                        acc[2] += sgn * d.z * dedr / r;
                        if (want_energy && j > i) E += e;
                    }
                }
            }
        }
        double* Fi = F + 3 * static_cast<size_t>(rows ? q : i);
        Fi[0] += acc[0]; Fi[1] += acc[1]; Fi[2] += acc[2];
    }
    return E;
}

}  // namespace

extern "C" {

// BAOAB like oracle_baoab with cell-list nonbonded forces on `nthreads` threads.
int oracle_baoab_cells(const ocm_system_desc* s, const ocm_integrator_desc* it, double* x, double* v,
                       int64_t step0, int64_t nsteps, int32_t nthreads) {
    const int n = s->n_particles;
    const Box box = make_box(s);
    std::vector<CellPairs> cps;
    for (int fi = 0; fi < s->n_forces; ++fi)
        if (!is_bonded(s->forces[fi].kind)) cps.push_back(make_cell_pairs(s->forces[fi], n));
    auto forces = [&](std::vector<double>& F) {
        std::fill(F.begin(), F.end(), 0.0);
        for (int fi = 0; fi < s->n_forces; ++fi)
            if (is_bonded(s->forces[fi].kind)) eval_bonded(s->forces[fi], x, box, F.data());
        for (const auto& cp : cps) pair_cells(s, cp, x, F.data(), nullptr, 0, nthreads, false);
    };
    const double kB = 1.380649e-23 * 6.02214076e23 / 1000.0;
    const double kT = kB * it->temperature, h = 0.5 * it->dt;
    const double c1 = std::exp(-it->friction * it->dt), c2 = 1.0 - c1 * c1;
    std::vector<double> F(3 * n), im(n);
    for (int i = 0; i < n; ++i) im[i] = s->masses[i] > 0 ? 1.0 / s->masses[i] : 0.0;
    forces(F);
    for (int64_t st = 0; st < nsteps; ++st) {
        const uint64_t sidx = static_cast<uint64_t>(step0 + st);
#pragma omp parallel for num_threads(nthreads)
        for (int i = 0; i < n; ++i) {
            if (im[i] == 0.0) continue;
            double g[3];
            gauss3(it->seed, sidx, static_cast<uint32_t>(i), 0u, g);
            const double sig = std::sqrt(kT * c2 * im[i]);
            for (int c = 0; c < 3; ++c) {
                double vv = v[3 * i + c] + h * F[3 * i + c] * im[i];
                double xx = x[3 * i + c] + h * vv;
                vv = c1 * vv + sig * g[c];
                xx = xx + h * vv;
                v[3 * i + c] = vv;
                x[3 * i + c] = xx;
            }
        }
        forces(F);
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) v[3 * i + c] += h * F[3 * i + c] * im[i];
    }
    return 0;
}

// oracle_eval through cell lists: forces of the groups in `groups` on every
// bead (rows == NULL; F n*3) or on beads rows[0..m) (F m*3, in that order),
// and, when E_by_group is given (rows must be NULL), the energy per group.
// F may be NULL.  OpenMP over beads for the nonbonded forces.
int oracle_eval_cells(const ocm_system_desc* s, const double* x, uint32_t groups, const int32_t* rows, int32_t m,
                      int32_t nthreads, double* F, double* E_by_group) {
    const int n = s->n_particles;
    const Box box = make_box(s);
    if (rows && E_by_group) return -1;
    const size_t nf = 3 * static_cast<size_t>(rows ? m : n);
    if (F) std::memset(F, 0, sizeof(double) * nf);
    if (E_by_group) std::memset(E_by_group, 0, sizeof(double) * 32);
    std::vector<double> Fb(3 * static_cast<size_t>(n), 0.0), Fp(nf, 0.0);
    for (int fi = 0; fi < s->n_forces; ++fi) {
        const ocm_force_desc& f = s->forces[fi];
        if (!((groups >> f.group) & 1u)) continue;
        double e;
        if (is_bonded(f.kind)) {
            e = eval_bonded(f, x, box, Fb.data());
        } else {
            const CellPairs cp = make_cell_pairs(f, n);
            e = pair_cells(s, cp, x, Fp.data(), rows, m, nthreads, E_by_group != nullptr);
        }
        if (E_by_group) E_by_group[f.group] += e;
    }
    if (F) {
        for (size_t q = 0; q < nf / 3; ++q) {
            const size_t b = rows ? static_cast<size_t>(rows[q]) : q;
            for (int c = 0; c < 3; ++c) F[3 * q + c] = Fb[3 * b + c] + Fp[3 * q + c];
        }
    }
    return 0;
}

}  // extern "C"
