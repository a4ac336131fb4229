// Host-side transform plan of the fast spectral operator (no HIP calls).
//
// The reference evaluates, per DOF and per (radial, direction) pair, two
// phase-shifted inverse 3D FFTs, a product and a forward FFT
// (frfs/solvers/dgfs/scattering.py:239-281; frfs/solvers/dgfsbi/scattering.py:274-315).
// This plan rewrites that work exactly (up to rounding) as
//
//   items    E_i = DFT+( phi(theta_i) . F_src(i) ),  phi(theta)(l) = prod_d e^{i theta_d l_d}
//   products P_k = coef_k * E_i1(k) . E_i2(k)             (pointwise, v space)
//   groups   G_g = sum_{k in g} P_k,  Ghat_g = DFT-(G_g)
//   contribs H_t += W_c . Ghat_g(c)                        (FP64, frequency space)
//   loss     L_t += Lt_j . F_src(j)
//   output   Q_t  = Re DFT+(H_t) - f_t . Re DFT+(L_t)
//
// using three exact identities:
//   * cos(a)F, sin(a)F -> E+ = DFT+(e^{ia}F), E- = DFT+(e^{-ia}F), and
//     A^2 + B^2 = E+ E-  (vhs-gll.mako:103-169);
//   * e^{i c sigma.l} is separable per axis, so no a-table is needed;
//   * identical (source, theta) inverse transforms and identical products
//     are computed once (antipodal design points make VSS same-species
//     products pairwise equal; cross-species pairs (p,q),(q,p) share too),
//     and for VHS the forward FFT is hoisted out of the direction sum
//     because bb1 depends on the radial index only (vhs-gll.mako:192-203).
#pragma once
#include <cmath>
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/frfs_fq.h"

namespace fq {

struct PItem { int src; double th[3]; };
struct PProd { int i1, i2; double coef; int group; };
struct PContrib { int group; int tgt; };
struct PLoss { int src; int tgt; };

// One term of a weight table (index `table` in [0, n_contribs + n_loss)).
//   kind 0 : scale * A * sinc(B |l|)                        (vhs bb1 / bb2 term)
//   kind 1 : scale * b_pq[x,y](l)   (vss-gll.mako:79-90, b part)
//   kind 2 : scale * c_pq(l)        (vss-gll.mako:79-92, c part, all x,y)
struct PTerm {
    int table;
    int kind;
    double A, B;
    int x, y, pair;
    double scale;
};

struct Plan {
    int model = 0, N = 0, Nrho = 0, M = 0, nspcs = 1, Mpre = 0;
    int nsrc = 1, ntgt = 1;
    double L = 0, sw = 0, cw = 0, swpre = 0;
    std::vector<double> qz, qw, sz, szpre, masses, eta, gam, prefac;
    std::vector<PItem> items;
    std::vector<PProd> prods;      // sorted by group
    int ngroups = 0;
    std::vector<int> group_p0, group_np;
    std::vector<PContrib> contribs;  // sorted by target
    std::vector<PLoss> loss;         // sorted by target
    std::vector<PTerm> terms;        // sorted by table
    int n_fft_ref = 0;

    int n_tables() const { return int(contribs.size() + loss.size()); }
};

inline bool plan_build(const fq_params &p, Plan &P, std::string &err) {
    if (p.abi_version != FQ_ABI_VERSION) { err = "fq_params.abi_version mismatch"; return false; }
    if (p.Nv <= 0 || p.Nv % 2) { err = "Nv must be positive and even"; return false; }
    if (p.Nrho <= 0 || p.M <= 0) { err = "Nrho and M must be positive"; return false; }
    if (!p.qz || !p.qw || !p.sz) { err = "qz, qw and sz are required"; return false; }
    if (!(p.L > 0)) { err = "L must be positive"; return false; }
    P.model = p.model; P.N = p.Nv; P.Nrho = p.Nrho; P.M = p.M;
    P.L = p.L; P.sw = p.sw; P.cw = p.cw;
    P.qz.assign(p.qz, p.qz + p.Nrho);
    P.qw.assign(p.qw, p.qw + p.Nrho);
    P.sz.assign(p.sz, p.sz + 3 * p.M);
    const int N = p.Nv;
    const double n = double(N) * N * N;

    std::map<std::tuple<int, double, double, double>, int> item_ix;
    auto item = [&](int src, double t0, double t1, double t2) {
        auto key = std::make_tuple(src, t0, t1, t2);
        auto it = item_ix.find(key);
        if (it != item_ix.end()) return it->second;
        int id = int(P.items.size());
        P.items.push_back({src, {t0, t1, t2}});
        item_ix[key] = id;
        return id;
    };
    std::vector<PProd> prods;
    std::vector<PTerm> terms;

    if (p.model == FQ_MODEL_VHS || p.model == FQ_MODEL_VHS_NODAL) {
        // frfs/solvers/dgfs/scattering.py:50-65,152-176; vhs-gll.mako:45-101,172-207
        const bool nodal = p.model == FQ_MODEL_VHS_NODAL;
        P.nspcs = 1;
        P.nsrc = nodal ? 2 : 1;
        P.ntgt = 1;
        const double gam = p.gamma;
        P.ngroups = p.Nrho;
        for (int r = 0; r < p.Nrho; ++r) {
            const double c = M_PI / p.L * p.qz[r] / 2.;          // math.pi/L*qz[p]/2.
            for (int q = 0; q < p.M; ++q) {
                const double t0 = c * p.sz[3 * q], t1 = c * p.sz[3 * q + 1], t2 = c * p.sz[3 * q + 2];
                if (!nodal) {
                    int ip = item(0, t0, t1, t2), im = item(0, -t0, -t1, -t2);
                    prods.push_back({ip, im, 1.0, r});
                } else {
                    // A_f A_g + B_f B_g = (E+_f E-_g + E-_f E+_g)/2
                    int fp = item(0, t0, t1, t2), fm = item(0, -t0, -t1, -t2);
                    int gp = item(1, t0, t1, t2), gm = item(1, -t0, -t1, -t2);
                    prods.push_back({fp, gm, 0.5, r});
                    prods.push_back({fm, gp, 0.5, r});
                }
            }
            P.contribs.push_back({r, 0});
            PTerm t{};
            t.table = r; t.kind = 0;
            t.A = std::pow(p.qz[r], gam + 2.) * 4 * M_PI;
            t.B = M_PI / p.L * p.qz[r] / 2.;
            t.scale = p.prefac * (2. * p.qw[r] * p.sw / n);
            terms.push_back(t);
        }
        // loss: bb2 (.) F_g for nodal, bb2 (.) F otherwise
        P.loss.push_back({nodal ? 1 : 0, 0});
        for (int r = 0; r < p.Nrho; ++r) {
            PTerm t{};
            t.table = p.Nrho; t.kind = 0;
            t.A = p.qw[r] * std::pow(p.qz[r], gam + 2.) * 16 * M_PI * M_PI;
            t.B = M_PI / p.L * p.qz[r];
            t.scale = p.prefac;
            terms.push_back(t);
        }
        const int MN = p.M * p.Nrho;
        P.n_fft_ref = nodal ? (5 * MN + 4) : (3 * MN + 3);
    } else if (p.model == FQ_MODEL_VSS) {
        // frfs/solvers/dgfsbi/scattering.py:52-91,172-200,259-330; vss-gll.mako:31-301
        const int ns = p.nspcs;
        if (ns < 1 || ns > 4) { err = "nspcs must be in [1, 4]"; return false; }
        if (!p.masses || !p.eta || !p.gammapq || !p.prefacpq || !p.szpre || p.Mpre <= 0) {
            err = "VSS needs masses, eta, gammapq, prefacpq, szpre and Mpre"; return false;
        }
        P.nspcs = ns; P.nsrc = ns; P.ntgt = ns; P.Mpre = p.Mpre; P.swpre = p.swpre;
        P.masses.assign(p.masses, p.masses + ns);
        P.eta.assign(p.eta, p.eta + ns * ns);
        P.gam.assign(p.gammapq, p.gammapq + ns * ns);
        P.prefac.assign(p.prefacpq, p.prefacpq + ns * ns);
        P.szpre.assign(p.szpre, p.szpre + 3 * p.Mpre);
        // (1 + sigma.sigma')^eta needs a positive base for non-integer eta
        for (int pq = 0; pq < ns * ns; ++pq) {
            double eta = P.eta[pq];
            if (eta == std::floor(eta)) continue;
            for (int y = 0; y < p.M; ++y)
                for (int k = 0; k < p.Mpre; ++k) {
                    double base = 1 + p.sz[3 * y] * p.szpre[3 * k] + p.sz[3 * y + 1] * p.szpre[3 * k + 1]
                                  + p.sz[3 * y + 2] * p.szpre[3 * k + 2];
                    // a negative base gives NaN and a zero base with eta < 0 gives inf in
                    // the reference's pow (vss-gll.mako:80-82)
                    if (base < 0 || (base == 0 && eta < 0)) {
                        err = "VSS pre-integration design is incompatible with the directions "
                              "(1 + sigma.sigma' < 0, or = 0 with alpha < 1); choose another Mpre";
                        return false;
                    }
                }
        }
        std::map<std::pair<int, int>, int> prod_group;
        std::map<std::pair<int, int>, int> contrib_ix;   // (group, tgt) -> contrib
        std::vector<PContrib> contribs;
        std::vector<std::vector<PTerm>> cterms;
        for (int pp = 0; pp < ns; ++pp)
            for (int qq = 0; qq < ns; ++qq) {
                const int pair = pp * ns + qq;
                const double mp = p.masses[pp], mq = p.masses[qq];
                const double kq = mq / (mp + mq), kp = mp / (mp + mq);
                for (int x = 0; x < p.Nrho; ++x) {
                    const double c = M_PI / p.L * p.qz[x];        // math.pi/L*qz[x]
                    for (int y = 0; y < p.M; ++y) {
                        const double s0 = c * p.sz[3 * y], s1 = c * p.sz[3 * y + 1], s2 = c * p.sz[3 * y + 2];
                        int i1 = item(pp, s0 * kq, s1 * kq, s2 * kq);
                        int i2 = item(qq, -(s0 * kp), -(s1 * kp), -(s2 * kp));
                        auto key = std::make_pair(std::min(i1, i2), std::max(i1, i2));
                        int g;
                        auto it = prod_group.find(key);
                        if (it == prod_group.end()) {
                            g = int(prod_group.size());
                            prod_group[key] = g;
                            prods.push_back({i1, i2, 1.0, g});
                        } else {
                            g = it->second;
                        }
                        auto ck = std::make_pair(g, pp);
                        int ci;
                        auto cit = contrib_ix.find(ck);
                        if (cit == contrib_ix.end()) {
                            ci = int(contribs.size());
                            contrib_ix[ck] = ci;
                            contribs.push_back({g, pp});
                            cterms.emplace_back();
                        } else {
                            ci = cit->second;
                        }
                        PTerm t{};
                        t.kind = 1; t.x = x; t.y = y; t.pair = pair;
                        t.scale = p.prefacpq[pair] * (p.qw[x] * p.sw / n);
                        cterms[ci].push_back(t);
                    }
                }
            }
        P.ngroups = int(prod_group.size());
        // contribs sorted by target (stable), terms renumbered
        for (int t = 0; t < ns; ++t)
            for (size_t ci = 0; ci < contribs.size(); ++ci)
                if (contribs[ci].tgt == t) {
                    int tab = int(P.contribs.size());
                    P.contribs.push_back(contribs[ci]);
                    for (auto tm : cterms[ci]) { tm.table = tab; terms.push_back(tm); }
                }
        // loss tables: target p, source q, prefac_pq c_pq
        const int nc = int(P.contribs.size());
        for (int pp = 0; pp < ns; ++pp)
            for (int qq = 0; qq < ns; ++qq) {
                int tab = nc + int(P.loss.size());
                P.loss.push_back({qq, pp});
                PTerm t{};
                t.table = tab; t.kind = 2; t.pair = pp * ns + qq;
                t.scale = p.prefacpq[pp * ns + qq];
                terms.push_back(t);
            }
        P.n_fft_ref = ns * ns * (3 * p.M * p.Nrho + 4);
    } else {
        err = "unknown model";
        return false;
    }
    // products sorted by group
    P.group_p0.assign(P.ngroups, 0);
    P.group_np.assign(P.ngroups, 0);
    for (auto &pr : prods) P.group_np[pr.group]++;
    for (int g = 1; g < P.ngroups; ++g) P.group_p0[g] = P.group_p0[g - 1] + P.group_np[g - 1];
    P.prods.resize(prods.size());
    std::vector<int> fill(P.ngroups, 0);
    for (auto &pr : prods) P.prods[P.group_p0[pr.group] + fill[pr.group]++] = pr;
    P.terms = terms;
    return true;
}

inline void plan_info(const Plan &P, fq_plan_info &i) {
    i.n_items = int(P.items.size());
    i.n_products = int(P.prods.size());
    i.n_groups = P.ngroups;
    i.n_contribs = int(P.contribs.size());
    i.n_loss = int(P.loss.size());
    i.n_sources = P.nsrc;
    i.n_targets = P.ntgt;
    i.n_fft_ref = P.n_fft_ref;
    i.n_fft = P.nsrc + i.n_items + P.ngroups + 2 * P.ntgt;
}

}  // namespace fq
