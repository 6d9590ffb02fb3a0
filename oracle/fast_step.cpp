// fast_step.cpp -- the CPU step the benchmark times as its baseline.
//
// TEST INFRASTRUCTURE ONLY, like oracle.cpp: loaded by tests/ and by
// bench.py's cpu_baseline leg, never by the product.
//
// The same BAOAB step as oracle_baoab (the same Philox noise, FP64
// throughout), organised as an optimised CPU MD code would be, so the
// baseline measures a CPU implementation rather than the oracle's
// straightforward one:
//   - a Verlet list per group of nonbonded forces that share cutoff,
//     exclusions and participation (LJ repulsive + attractive in one pass),
//     built from a cell list (members counting-sorted by cell, a SIMD
//     distance test per stencil cell) with a skin, and rebuilt only when a
//     bead has moved more than skin / 2 since the build;
//   - both directions of every pair in the list, so each bead sums its own
//     force with no scatter: OpenMP over beads, 4-lane AVX2 SIMD with
//     gathers over a row (the row loop is branch-free and templated on the
//     group's kinds; the Makefile tunes for the EPYC hosts, which makes the
//     compiler emit the gathers);
//   - no pow(): powers by multiplication; exp by a vectorisable
//     Cody-Waite + Taylor routine;
//   - two-body bonded terms (harmonic bonds, Go contacts) gathered per bead
//     in parallel; other bonded kinds through oracle_eval.
// Kinds: every bonded kind; nonbonded excluded volume, Debye-Hueckel and LJ
// repulsive/attractive (the kinds of the benchmark systems).  Other
// nonbonded kinds return -2.
//
// Formulas follow oracle.cpp's pair_energy and eval_bonded (which cite the
// reference's generators); tests/test_oracle.py pins this step against
// oracle_baoab over 30 steps with list rebuilds.

#include "../include/ocm.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {
int oracle_eval(const ocm_system_desc* s, const double* x, uint32_t groups, double* F, double* E_by_group);
void oracle_gauss3(uint64_t seed, uint64_t step, uint32_t bead, uint32_t stream, double* g);
}

namespace {

// exp(a) for a <= 0 in a form the vectoriser keeps in SIMD registers (the
// library exp is a scalar call in a vector loop): Cody-Waite reduction
// a = k ln2 + r with ln 2 split as in fdlibm, exp(r) by its Taylor series to
// r^12 (|r| <= ln2/2: truncation < 2e-16 relative), 2^k assembled in the
// exponent field from the integer that the 1.5 * 2^52 shift leaves in the
// low mantissa bits.
inline double exp_nonpositive(double a) {
    a = std::max(a, -700.0);
    const double shift = 6755399441055744.0;   // 1.5 * 2^52
    const double t = std::fma(a, 1.4426950408889634, shift);
    const double k = t - shift;
    double r = std::fma(-k, 6.93147180369123816490e-01, a);
    r = std::fma(-k, 1.90821492927058770002e-10, r);
    double p = 1.0 / 479001600.0;   // 1/12!
    p = std::fma(p, r, 1.0 / 39916800.0);
    p = std::fma(p, r, 1.0 / 3628800.0);
    p = std::fma(p, r, 1.0 / 362880.0);
    p = std::fma(p, r, 1.0 / 40320.0);
    p = std::fma(p, r, 1.0 / 5040.0);
    p = std::fma(p, r, 1.0 / 720.0);
    p = std::fma(p, r, 1.0 / 120.0);
    p = std::fma(p, r, 1.0 / 24.0);
    p = std::fma(p, r, 1.0 / 6.0);
    p = std::fma(p, r, 0.5);
    p = std::fma(p, r, 1.0);
    p = std::fma(p, r, 1.0);
    uint64_t tb;
    std::memcpy(&tb, &t, sizeof tb);
    const uint64_t sb = (tb + 1023u) << 52;   // the low bits of t hold k (two's complement)
    double scale;
    std::memcpy(&scale, &sb, sizeof scale);
    return p * scale;
}

// Per-bead parameters of one group's kinds, de-interleaved.
struct GroupParams {
    std::vector<double> pe, pd, re, rs, ae, as;   // exv sigma, DH charge, LJ rep eps/sigma, LJ attr eps/sigma
};

struct PairGroupCpu {
    std::vector<const ocm_force_desc*> forces;
    bool periodic = false;
    double rc = 0, rc2 = 0;
    std::vector<int> members;                 // participating beads
    std::vector<int> ex_start, ex;            // excluded partners per bead, sorted (set up once)
    std::vector<int> row_start, col;          // Verlet list over members (both directions)
    GroupParams p;
};

bool same_list(const ocm_force_desc& a, const ocm_force_desc& b, int n) {
    if (a.use_periodic != b.use_periodic || a.cutoff != b.cutoff || a.n_exclusions != b.n_exclusions) return false;
    if (a.n_exclusions && std::memcmp(a.exclusions, b.exclusions, sizeof(int32_t) * 2 * a.n_exclusions) != 0)
        return false;
    if ((a.category == nullptr) != (b.category == nullptr)) return false;
    if (a.category) {
        if (a.n_categories != b.n_categories || std::memcmp(a.category, b.category, sizeof(int32_t) * n) != 0)
            return false;
        const bool aa = a.category_allow != nullptr, ba = b.category_allow != nullptr;
        if (aa != ba) return false;
        if (aa && std::memcmp(a.category_allow, b.category_allow, static_cast<size_t>(a.n_categories) * a.n_categories))
            return false;
    }
    return true;
}

struct Stepper {
    const ocm_system_desc* s = nullptr;
    int n = 0;
    double L[3] = {0, 0, 0}, iL[3] = {0, 0, 0};
    bool box = false;
    double skin = 0.3;
    int nthreads = 1;
    std::vector<PairGroupCpu> groups;
    std::vector<double> ref;                  // positions at the last build
    uint32_t other_groups = 0;                // groups evaluated through oracle_eval
    // two-body bonded terms gathered per bead: (term kind, partner, p0, p1)
    std::vector<int> b2_start, b2_kind, b2_partner;
    std::vector<double> b2_p0, b2_p1;
    std::vector<double> tmp;
    std::vector<double> xs, ys, zs;           // positions by axis, refreshed per force evaluation
    long long rebuilds = 0;

    double min_image(double d, int c, bool pbc) const {
        return (pbc && box) ? std::fma(-L[c], std::rint(d * iL[c]), d) : d;
    }

    int setup() {
        n = s->n_particles;
        box = s->use_periodic_box != 0;
        for (int c = 0; c < 3; ++c) {
            L[c] = s->box[c];
            iL[c] = box ? 1.0 / s->box[c] : 0.0;
        }
        std::vector<std::vector<int>> b2(n);
        std::vector<int> b2k, b2o;
        std::vector<double> p0, p1;
        for (int fi = 0; fi < s->n_forces; ++fi) {
            const ocm_force_desc& f = s->forces[fi];
            if (f.kind < OCM_EXCLUDED_VOLUME) {
                if (f.kind == OCM_HARMONIC_BOND || f.kind == OCM_GO_CONTACT) {
                    for (int t = 0; t < f.n_terms; ++t) {
                        const int i = f.indices[2 * t], j = f.indices[2 * t + 1];
                        for (int side = 0; side < 2; ++side) {
                            const int a = side ? j : i, o = side ? i : j;
                            b2[a].push_back(static_cast<int>(b2k.size()));
                            b2k.push_back(f.kind | (f.use_periodic ? 256 : 0));
                            b2o.push_back(o);
                            p0.push_back(f.params[2 * t]);
                            p1.push_back(f.params[2 * t + 1]);
                        }
                    }
                } else {
                    other_groups |= 1u << f.group;
                }
                continue;
            }
            if (f.kind != OCM_EXCLUDED_VOLUME && f.kind != OCM_DEBYE_HUCKEL && f.kind != OCM_LJ_REPULSIVE &&
                f.kind != OCM_LJ_ATTRACTIVE)
                return -2;
            bool placed = false;
            for (auto& g : groups) {
                bool kind_taken = false;
                for (const auto* h : g.forces) kind_taken = kind_taken || h->kind == f.kind;
                if (!kind_taken && same_list(*g.forces[0], f, n)) { g.forces.push_back(&f); placed = true; break; }
            }
            if (!placed) {
                PairGroupCpu g;
                g.forces.push_back(&f);
                g.periodic = f.use_periodic != 0 && box;
                g.rc = f.cutoff;
                g.rc2 = f.cutoff * f.cutoff;
                for (int i = 0; i < n; ++i)
                    if (!f.category || f.category[i] >= 0) g.members.push_back(i);
                groups.push_back(std::move(g));
            }
        }
        for (auto& g : groups) {
            const ocm_force_desc& f = *g.forces[0];
            std::vector<std::vector<int>> ex(n);
            for (int64_t e = 0; e < f.n_exclusions; ++e) {
                const int a = f.exclusions[2 * e], b = f.exclusions[2 * e + 1];
                ex[a].push_back(b);
                ex[b].push_back(a);
            }
            g.ex_start.assign(n + 1, 0);
            for (int i = 0; i < n; ++i) {
                std::sort(ex[i].begin(), ex[i].end());
                g.ex_start[i + 1] = g.ex_start[i] + static_cast<int>(ex[i].size());
                g.ex.insert(g.ex.end(), ex[i].begin(), ex[i].end());
            }
        }
        for (auto& g : groups) {
            for (const auto* f : g.forces) {
                const double* q = f->particle_params;
                auto take = [&](std::vector<double>& d, int stride, int off) {
                    d.resize(n);
                    for (int i = 0; i < n; ++i) d[i] = q[stride * i + off];
                };
                if (f->kind == OCM_EXCLUDED_VOLUME) take(g.p.pe, 1, 0);
                else if (f->kind == OCM_DEBYE_HUCKEL) take(g.p.pd, 1, 0);
                else if (f->kind == OCM_LJ_REPULSIVE) { take(g.p.re, 2, 0); take(g.p.rs, 2, 1); }
                else { take(g.p.ae, 2, 0); take(g.p.as, 2, 1); }
            }
        }
        xs.assign(n, 0.0);
        ys.assign(n, 0.0);
        zs.assign(n, 0.0);
        b2_start.assign(n + 1, 0);
        for (int i = 0; i < n; ++i) b2_start[i + 1] = b2_start[i] + static_cast<int>(b2[i].size());
        for (int i = 0; i < n; ++i)
            for (int k : b2[i]) {
                b2_kind.push_back(b2k[k]);
                b2_partner.push_back(b2o[k]);
                b2_p0.push_back(p0[k]);
                b2_p1.push_back(p1[k]);
            }
        tmp.assign(3 * static_cast<size_t>(n), 0.0);
        return 0;
    }

    // Verlet lists of every group at positions x, with radius cutoff + skin:
    // members sorted by cell (counting sort), each cell's positions
    // contiguous and wrapped into the box, the distance test over a stencil
    // cell as one SIMD loop, the passing candidates compacted after it.
    void build(const double* x) {
        ++rebuilds;
        ref.assign(x, x + 3 * static_cast<size_t>(n));
        for (auto& g : groups) {
            const ocm_force_desc& f = *g.forces[0];
            const double rl = g.rc + skin, rl2 = rl * rl;
            double lo[3], ext[3];
            int nc[3];
            for (int c = 0; c < 3; ++c) {
                if (g.periodic) {
                    lo[c] = 0.0;
                    ext[c] = L[c];
                } else {
                    double mn = 1e300, mx = -1e300;
                    for (int i : g.members) { mn = std::min(mn, x[3 * i + c]); mx = std::max(mx, x[3 * i + c]); }
                    lo[c] = mn;
                    ext[c] = std::max(mx - mn, 1e-9) * (1.0 + 1e-12);
                }
                nc[c] = std::max(1, std::min(static_cast<int>(ext[c] / rl), 512));
                if (g.periodic && nc[c] < 3) nc[c] = 1;
            }
            const int m = static_cast<int>(g.members.size());
            const int ncell = nc[0] * nc[1] * nc[2];
            std::vector<int> cell_of(m), cstart(ncell + 1, 0);
            std::vector<double> wx(m), wy(m), wz(m);
            for (int q = 0; q < m; ++q) {
                const int i = g.members[q];
                int ci[3];
                double w[3];
                for (int c = 0; c < 3; ++c) {
                    w[c] = x[3 * i + c];
                    if (g.periodic) w[c] -= L[c] * std::floor(w[c] * iL[c]);
                    const double u = (w[c] - lo[c]) / ext[c];
                    ci[c] = std::max(0, std::min(nc[c] - 1, static_cast<int>(u * nc[c])));
                }
                wx[q] = w[0]; wy[q] = w[1]; wz[q] = w[2];
                cell_of[q] = (ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
                ++cstart[cell_of[q] + 1];
            }
            for (int c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
            std::vector<int> fill(cstart.begin(), cstart.end() - 1), sorted(m);
            std::vector<double> sx(m), sy(m), sz(m);
            for (int q = 0; q < m; ++q) {
                const int at = fill[cell_of[q]]++;
                sorted[at] = q;
                sx[at] = wx[q]; sy[at] = wy[q]; sz[at] = wz[q];
            }
            std::vector<std::vector<int>> rows(m);
#pragma omp parallel num_threads(nthreads)
            {
                std::vector<unsigned char> pass;
#pragma omp for schedule(dynamic, 64)
                for (int q = 0; q < m; ++q) {
                    const int i = g.members[q];
                    const double xi = wx[q], yi = wy[q], zi = wz[q];
                    const int cell = cell_of[q];
                    const int cx = cell % nc[0], cy = (cell / nc[0]) % nc[1], cz = cell / (nc[0] * nc[1]);
                    auto& row = rows[q];
                    const int eb = g.ex_start[i], ee = g.ex_start[i + 1];
                    for (int dz = -1; dz <= 1; ++dz) {
                        if (nc[2] == 1 && dz) continue;
                        int z = cz + dz;
                        double shz = 0.0;
                        if (g.periodic) { if (z < 0) { z += nc[2]; shz = -L[2]; } else if (z >= nc[2]) { z -= nc[2]; shz = L[2]; } }
                        else if (z < 0 || z >= nc[2]) continue;
                        for (int dy = -1; dy <= 1; ++dy) {
                            if (nc[1] == 1 && dy) continue;
                            int y = cy + dy;
                            double shy = 0.0;
                            if (g.periodic) { if (y < 0) { y += nc[1]; shy = -L[1]; } else if (y >= nc[1]) { y -= nc[1]; shy = L[1]; } }
                            else if (y < 0 || y >= nc[1]) continue;
                            for (int dx = -1; dx <= 1; ++dx) {
                                if (nc[0] == 1 && dx) continue;
                                int xx = cx + dx;
                                double shx = 0.0;
                                if (g.periodic) { if (xx < 0) { xx += nc[0]; shx = -L[0]; } else if (xx >= nc[0]) { xx -= nc[0]; shx = L[0]; } }
                                else if (xx < 0 || xx >= nc[0]) continue;
                                const int c = (z * nc[1] + y) * nc[0] + xx;
                                const int k0 = cstart[c], k1 = cstart[c + 1];
                                pass.resize(static_cast<size_t>(k1 - k0));
                                // a periodic axis with one cell takes the minimum image per pair
                                const bool mi0 = g.periodic && nc[0] == 1, mi1 = g.periodic && nc[1] == 1,
                                           mi2 = g.periodic && nc[2] == 1;
#pragma omp simd
                                for (int k = k0; k < k1; ++k) {
                                    double ddx = sx[k] + shx - xi, ddy = sy[k] + shy - yi, ddz = sz[k] + shz - zi;
                                    if (mi0) ddx -= L[0] * std::rint(ddx * iL[0]);
                                    if (mi1) ddy -= L[1] * std::rint(ddy * iL[1]);
                                    if (mi2) ddz -= L[2] * std::rint(ddz * iL[2]);
                                    pass[k - k0] = ddx * ddx + ddy * ddy + ddz * ddz < rl2;
                                }
                                for (int k = k0; k < k1; ++k) {
                                    if (!pass[k - k0]) continue;
                                    const int j = g.members[sorted[k]];
                                    if (j == i) continue;
                                    if (f.category && f.category_allow &&
                                        !f.category_allow[f.category[i] * f.n_categories + f.category[j]])
                                        continue;
                                    if (ee > eb && std::binary_search(g.ex.begin() + eb, g.ex.begin() + ee, j)) continue;
                                    row.push_back(j);
                                }
                            }
                        }
                    }
                }
            }
            g.row_start.assign(m + 1, 0);
            for (int q = 0; q < m; ++q) g.row_start[q + 1] = g.row_start[q] + static_cast<int>(rows[q].size());
            g.col.resize(g.row_start[m]);
#pragma omp parallel for num_threads(nthreads)
            for (int q = 0; q < m; ++q) std::copy(rows[q].begin(), rows[q].end(), g.col.begin() + g.row_start[q]);
        }
    }

    bool list_stale(const double* x) const {
        const double h2 = 0.25 * skin * skin;
        bool stale = false;
#pragma omp parallel for num_threads(nthreads) reduction(|| : stale)
        for (int i = 0; i < n; ++i) {
            const double dx = x[3 * i] - ref[3 * i], dy = x[3 * i + 1] - ref[3 * i + 1], dz = x[3 * i + 2] - ref[3 * i + 2];
            stale = stale || !(dx * dx + dy * dy + dz * dz <= h2);
        }
        return stale;
    }

    void forces(const double* x, double* F) {
        std::fill(F, F + 3 * static_cast<size_t>(n), 0.0);
        if (other_groups) {
            oracle_eval(s, x, other_groups, tmp.data(), nullptr);
            for (size_t k = 0; k < tmp.size(); ++k) F[k] += tmp[k];
        }
        // two-body bonded terms, each bead summing its own
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int i = 0; i < n; ++i) {
            double fx = 0, fy = 0, fz = 0;
            for (int k = b2_start[i]; k < b2_start[i + 1]; ++k) {
                const int j = b2_partner[k];
                const bool pbc = (b2_kind[k] & 256) != 0;
                const double dx = min_image(x[3 * j] - x[3 * i], 0, pbc);
                const double dy = min_image(x[3 * j + 1] - x[3 * i + 1], 1, pbc);
                const double dz = min_image(x[3 * j + 2] - x[3 * i + 2], 2, pbc);
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                double dedr;
                if ((b2_kind[k] & 255) == OCM_HARMONIC_BOND) {       // {r0, k}
                    dedr = b2_p1[k] * (r - b2_p0[k]);
                } else {                                            // Go contact {k, r0}
                    const double q = b2_p1[k] / r, q2 = q * q, q4 = q2 * q2, q10 = q4 * q4 * q2, q12 = q10 * q2;
                    dedr = 60.0 * b2_p0[k] * (q10 - q12) / r;
                }
                const double g = dedr / r;   // force on i is +g d (d = x_j - x_i)
                fx += g * dx; fy += g * dy; fz += g * dz;
            }
            F[3 * i] += fx; F[3 * i + 1] += fy; F[3 * i + 2] += fz;
        }
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int i = 0; i < n; ++i) {
            xs[i] = x[3 * i];
            ys[i] = x[3 * i + 1];
            zs[i] = x[3 * i + 2];
        }
        for (const auto& g : groups) pair_forces(g, F);
    }

    // The kinds of one group, their constants, and the row loop over its
    // list written branch-free (selects instead of branches, positions and
    // parameters by bead in separate arrays) so that the compiler turns it
    // into 4-lane SIMD with gathers.
    struct PairConsts {
        const double *pe = nullptr, *pd = nullptr, *re = nullptr, *rs = nullptr, *ae = nullptr, *as = nullptr;
        double exv_eps = 0, dh_K = 0, dh_il = 0, rsw = 0, isw = 0, rc2 = 0;
    };

    template <bool EXV, bool DH, bool LJR, bool LJA, bool SW, bool PBC>
    void pair_rows(const PairGroupCpu& g, const PairConsts& k, const double* __restrict X, const double* __restrict Y,
                   const double* __restrict Z, double* F) {
        const double two6sq = 1.2599210498948732;   // 2^(1/3) = (2^(1/6))^2
        const double Lx = L[0], Ly = L[1], Lz = L[2], iLx = iL[0], iLy = iL[1], iLz = iL[2];
        const int* __restrict col = g.col.data();
        const int m = static_cast<int>(g.members.size());
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 32)
        for (int q = 0; q < m; ++q) {
            const int i = g.members[q];
            const double xi = X[i], yi = Y[i], zi = Z[i];
            const double pei = EXV ? k.pe[i] : 0.0, pdi = DH ? k.pd[i] : 0.0;
            const double eri = LJR ? k.re[i] : 0.0, sri = LJR ? k.rs[i] : 0.0;
            const double eai = LJA ? k.ae[i] : 0.0, sai = LJA ? k.as[i] : 0.0;
            double fx = 0, fy = 0, fz = 0;
            const int e0 = g.row_start[q], e1 = g.row_start[q + 1];
#pragma omp simd reduction(+ : fx, fy, fz)
            for (int e = e0; e < e1; ++e) {
                const int j = col[e];
                double dx = X[j] - xi, dy = Y[j] - yi, dz = Z[j] - zi;
                if (PBC) {
                    dx -= Lx * std::rint(dx * iLx);
                    dy -= Ly * std::rint(dy * iLy);
                    dz -= Lz * std::rint(dz * iLz);
                }
                const double r2raw = dx * dx + dy * dy + dz * dz;
                const bool in = r2raw < k.rc2;
                const double r2 = in ? r2raw : 1.0;
                const double ir2 = 1.0 / r2;
                double gsum = 0;   // (dE/dr) / r; the force on i is +gsum d
                if (EXV) {
                    const double sg = pei + k.pe[j];
                    const double q2 = sg * sg * ir2, q6 = q2 * q2 * q2;
                    gsum += -12.0 * k.exv_eps * q6 * q6 * ir2;
                }
                if (DH) {
                    const double r = std::sqrt(r2), ir = 1.0 / r;
                    gsum += -pdi * k.pd[j] * k.dh_K * exp_nonpositive(-r * k.dh_il) * (ir2 + k.dh_il * ir) * ir;
                }
                if (LJR) {   // the repulsive core up to r* = 2^(1/6) sigma
                    const double sb = 0.5 * (sri + k.rs[j]);
                    const double s2 = sb * sb * ir2, s6 = s2 * s2 * s2;
                    const double v = 0.5 * (eri + k.re[j]) * 4.0 * (-12.0 * s6 * s6 + 6.0 * s6) * ir2;
                    gsum += (sb * sb * two6sq - r2 >= 0.0) ? v : 0.0;
                }
                if (LJA) {   // the attractive well beyond r*, -epsilon inside it, switched to zero at rc
                    const double sb = 0.5 * (sai + k.as[j]);
                    const double s2 = sb * sb * ir2, s6 = s2 * s2 * s2, s12 = s6 * s6;
                    const double inside = sb * sb * two6sq - r2;
                    const double ea = 0.5 * (eai + k.ae[j]);
                    const double outer = inside <= 0.0 ? 1.0 : 0.0;
                    const double u = outer * ea * 4.0 * (s12 - s6) - (inside >= 0.0 ? ea : 0.0);
                    double du_r = outer * ea * 4.0 * (-12.0 * s12 + 6.0 * s6) * ir2;
                    if (SW) {
                        const double r = std::sqrt(r2);
                        const double t = std::max(r - k.rsw, 0.0) * k.isw, t2 = t * t, t3 = t2 * t;
                        const double S = 1.0 - t3 * (10.0 - 15.0 * t + 6.0 * t2);
                        const double dS = -30.0 * t2 * (1.0 - t) * (1.0 - t) * k.isw;
                        du_r = du_r * S + u * dS / r;
                    }
                    gsum += du_r;
                }
                gsum = in ? gsum : 0.0;
                fx += gsum * dx;
                fy += gsum * dy;
                fz += gsum * dz;
            }
            F[3 * i] += fx;
            F[3 * i + 1] += fy;
            F[3 * i + 2] += fz;
        }
    }

    template <bool EXV, bool DH, bool LJR, bool LJA, bool SW>
    void pair_rows_pbc(const PairGroupCpu& g, const PairConsts& k, const double* X, const double* Y, const double* Z,
                       double* F) {
        if (g.periodic) pair_rows<EXV, DH, LJR, LJA, SW, true>(g, k, X, Y, Z, F);
        else pair_rows<EXV, DH, LJR, LJA, SW, false>(g, k, X, Y, Z, F);
    }

    void pair_forces(const PairGroupCpu& g, double* F) {
        const GroupParams& p = g.p;
        const double *X = xs.data(), *Y = ys.data(), *Z = zs.data();
        const ocm_force_desc *exv = nullptr, *dh = nullptr, *ljr = nullptr, *lja = nullptr;
        for (const auto* f : g.forces) {
            if (f->kind == OCM_EXCLUDED_VOLUME) exv = f;
            else if (f->kind == OCM_DEBYE_HUCKEL) dh = f;
            else if (f->kind == OCM_LJ_REPULSIVE) ljr = f;
            else lja = f;
        }
        PairConsts k;
        k.rc2 = g.rc2;
        if (exv) { k.pe = p.pe.data(); k.exv_eps = exv->globals[0]; }
        if (dh) { k.pd = p.pd.data(); k.dh_K = dh->globals[0]; k.dh_il = 1.0 / dh->globals[1]; }
        if (ljr) { k.re = p.re.data(); k.rs = p.rs.data(); }
        bool sw = false;
        if (lja) {
            k.ae = p.ae.data();
            k.as = p.as.data();
            const double rsw = lja->switch_distance;
            sw = rsw > 0.0;
            k.rsw = rsw;
            k.isw = sw ? 1.0 / (g.rc - rsw) : 0.0;
        }
        // the compositions the groups take: one kind, or LJ repulsive + attractive
        if (exv && !dh && !ljr && !lja) pair_rows_pbc<true, false, false, false, false>(g, k, X, Y, Z, F);
        else if (dh && !exv && !ljr && !lja) pair_rows_pbc<false, true, false, false, false>(g, k, X, Y, Z, F);
        else if (ljr && lja && !exv && !dh && sw) pair_rows_pbc<false, false, true, true, true>(g, k, X, Y, Z, F);
        else if (ljr && lja && !exv && !dh) pair_rows_pbc<false, false, true, true, false>(g, k, X, Y, Z, F);
        else if (ljr && !lja && !exv && !dh) pair_rows_pbc<false, false, true, false, false>(g, k, X, Y, Z, F);
        else if (lja && !ljr && !exv && !dh && sw) pair_rows_pbc<false, false, false, true, true>(g, k, X, Y, Z, F);
        else if (lja && !ljr && !exv && !dh) pair_rows_pbc<false, false, false, true, false>(g, k, X, Y, Z, F);
        else if (sw) pair_rows_pbc<true, true, true, true, true>(g, k, X, Y, Z, F);   // any other mix
        else pair_rows_pbc<true, true, true, true, false>(g, k, X, Y, Z, F);
    }

};

}  // namespace

extern "C" {

// Forces of every group at x through the fast step's lists (built at x).
int oracle_fast_forces(const ocm_system_desc* s, const double* x, double skin, int32_t nthreads, double* F) {
    Stepper st;
    st.s = s;
    st.skin = skin;
    st.nthreads = nthreads;
    const int rc = st.setup();
    if (rc) return rc;
    st.build(x);
    st.forces(x, F);
    return 0;
}

// BAOAB as oracle_baoab (same splitting, same Philox stream 0 noise), with
// the fast force evaluation; *rebuilds receives the list builds done.
int oracle_baoab_fast(const ocm_system_desc* s, const ocm_integrator_desc* it, double* x, double* v,
                      int64_t step0, int64_t nsteps, int32_t nthreads, double skin, int64_t* rebuilds) {
    Stepper st;
    st.s = s;
    st.skin = skin;
    st.nthreads = nthreads;
    const int rc = st.setup();
    if (rc) return rc;
    const int n = s->n_particles;
    const double kB = 1.380649e-23 * 6.02214076e23 / 1000.0;
    const double kT = kB * it->temperature, h = 0.5 * it->dt;
    const double c1 = std::exp(-it->friction * it->dt), c2 = 1.0 - c1 * c1;
    std::vector<double> F(3 * static_cast<size_t>(n)), im(n);
    for (int i = 0; i < n; ++i) im[i] = s->masses[i] > 0 ? 1.0 / s->masses[i] : 0.0;
    st.build(x);
    st.forces(x, F.data());
    for (int64_t t = 0; t < nsteps; ++t) {
        const uint64_t sidx = static_cast<uint64_t>(step0 + t);
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int i = 0; i < n; ++i) {
            if (im[i] == 0.0) continue;
            double g[3];
            oracle_gauss3(it->seed, sidx, static_cast<uint32_t>(i), 0u, g);
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
        if (st.list_stale(x)) st.build(x);
        st.forces(x, F.data());
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) v[3 * i + c] += h * F[3 * i + c] * im[i];
    }
    if (rebuilds) *rebuilds = st.rebuilds;
    return 0;
}

}  // extern "C"
