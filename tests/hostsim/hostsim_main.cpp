// Drives the product's host runtime (wx_host.cpp built against the host-only
// HIP stand-in) through many admesh rounds with seeded refinement flags, and
// checks its tree (lp, leaf, icoord, nbmax, nbleafs) against the oracle's
// admesh given the same flags after every round.  Built with ASan/UBSan.
// Usage: hostsim_main <rounds> <seed> [ens].  With "ens", three instances of
// different root layouts, depths and capacities run as one wx_ens ensemble
// and each instance's tree is compared with its own oracle mesh.
// TEST SCAFFOLDING.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../include/wx_capi.h"
#include "../../oracle/wxoracle.h"

extern "C" void wxsim_set_flags(const signed char *f, int n);
extern "C" int wxsim_ops(int *out, int cap);
extern "C" void wxsim_set_group_flags(int group, const signed char *f, int n);

static int run_ensemble(int rounds, unsigned seed, int groups);

int main(int argc, char **argv) {
  const int rounds = argc > 1 ? atoi(argv[1]) : 200;
  const unsigned seed = argc > 2 ? (unsigned)atoi(argv[2]) : 1u;
  if (argc > 3 && std::string(argv[3]) == "ens") return run_ensemble(rounds, seed, argc > 4 ? atoi(argv[4]) : 1);
  const int cap = 30000, nrx = 8, nry = 8, nlevs = 5, neq = 4;
  wx_params p{};
  p.neq = neq; p.nlevs = nlevs; p.nbrootx = nrx; p.nbrooty = nry; p.nblocks = cap;
  p.cv = 1.5; p.gamma = 5. / 3.; p.cfl = 0.4; p.eta = 1e-3; p.tol = 0.05; p.tol2 = 1.0;
  p.compat = 3; p.Tempsc = 1; p.Psc = 1; p.tsc = 1;
  for (int l = 0; l < nlevs; ++l) { p.dx[l] = 1.0 / (16 * nrx) / (1 << l); p.dy[l] = 1.0 / (16 * nry) / (1 << l); }
  wx_ctx *c = nullptr;
  if (wx_create(&p, &c)) { fprintf(stderr, "create failed\n"); return 2; }
  std::vector<double> u((size_t)(nrx * nry + 1) * neq * 400, 1.0);
  if (wx_basegrid(c, u.data(), nrx * nry)) { fprintf(stderr, "basegrid failed\n"); return 2; }
  // oracle mesh
  std::vector<double> ou((size_t)(cap + 1) * neq * 400), oup(ou.size()), opr(ou.size());
  std::vector<int32_t> olp((size_t)(cap + 1) * 8), oleaf(cap + 1), oic((size_t)(cap + 1) * 2), ob((size_t)(cap + 1) * 12);
  wxo_mesh m{cap, 0, 0, 0, ou.data(), oup.data(), opr.data(), olp.data(), oleaf.data(), oic.data(), ob.data()};
  wxo_basegrid((const wxo_params *)&p, &m);
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> U01(0.0, 1.0);
  std::vector<int32_t> flp, fleaf, fic;
  int total_ops = 0;
  for (int r = 0; r < rounds; ++r) {
    std::vector<signed char> f(cap + 1, 0);
    std::vector<int32_t> up(cap + 1, 0), dn(cap + 1, 0);
    const double pup = (r % 40 < 20) ? 0.06 : 0.01, pdn = (r % 40 < 20) ? 0.3 : 0.8;
    for (int nb = 1; nb <= m.nbmax; ++nb) {
      if (!oleaf[nb]) continue;
      const int lev = olp[(size_t)nb * 8];
      const double x = U01(rng);
      if (x < pup && lev < nlevs) { f[nb] = 1; up[nb] = 1; }
      else if (x < pup + pdn) { f[nb] = 2; dn[nb] = 1; }
    }
    wxsim_set_flags(f.data(), cap + 1);
    double t, dt, z;
    if (wx_step(c, 1, 1, &t, &dt, &z)) {
      char b[512];
      wx_last_error(b, 512);
      fprintf(stderr, "round %d: wx_step failed: %s\n", r, b);
      return 3;
    }
    int oerr = wxo_admesh_flags((const wxo_params *)&p, &m, up.data(), dn.data());
    if (oerr) { fprintf(stderr, "round %d: oracle admesh error %d\n", r, oerr); return 3; }
    std::vector<int> ops(6 * 20000);
    total_ops += wxsim_ops(ops.data(), 20000);
    int32_t nbmax = 0, nbleafs = 0;
    const int ld = m.nbmax;
    flp.assign((size_t)8 * ld, 0); fleaf.assign(ld, 0); fic.assign((size_t)2 * ld, 0);
    if (wx_get_tree(c, flp.data(), fleaf.data(), fic.data(), ld, &nbmax, &nbleafs)) { fprintf(stderr, "get_tree\n"); return 3; }
    if (nbmax != m.nbmax || nbleafs != m.nbleafs) {
      fprintf(stderr, "round %d: nbmax %d/%d nbleafs %d/%d\n", r, nbmax, m.nbmax, nbleafs, m.nbleafs);
      return 1;
    }
    for (int nb = 1; nb <= nbmax; ++nb) {
      for (int k = 1; k <= 8; ++k)
        if (flp[(size_t)(k - 1) * ld + nb - 1] != olp[(size_t)nb * 8 + k - 1]) {
          fprintf(stderr, "round %d: lp(%d,%d) %d vs oracle %d\n", r, nb, k, flp[(size_t)(k - 1) * ld + nb - 1], olp[(size_t)nb * 8 + k - 1]);
          return 1;
        }
      if (fleaf[nb - 1] != oleaf[nb] || fic[nb - 1] != oic[(size_t)nb * 2] || fic[(size_t)ld + nb - 1] != oic[(size_t)nb * 2 + 1]) {
        fprintf(stderr, "round %d: leaf/icoord of %d differ\n", r, nb);
        return 1;
      }
    }
    int32_t nint = 0;
    if (wx_locatebounds(c, &nint)) { fprintf(stderr, "locatebounds failed\n"); return 3; }
    if (nint != wxo_locatebounds((const wxo_params *)&p, &m)) { fprintf(stderr, "round %d: list length differs\n", r); return 1; }
  }
  printf("hostsim: %d rounds, %d mesh ops, final nbmax %d leaves %d -- trees identical\n", rounds, total_ops, m.nbmax, m.nbleafs);
  wx_destroy(c);
  return 0;
}

// ---------------------------------------------------------------------
// Ensemble mode: the instances are split into `groups` groups of
// consecutive instances (n / groups each, the first n % groups one more);
// within its group, instance i's block ID nb sits at slot base_i + nb with
// base_i = sum over the group's earlier instances j of (nblocks_j + 1).
// Each group reads its flags from a table of its own (jet[11] = group).
struct OMesh {
  wx_params p{};
  std::vector<double> u, up, pr;
  std::vector<int32_t> lp, leaf, ic, b;
  wxo_mesh m{};
  void init(const wx_params &pp) {
    p = pp;
    const int cap = pp.nblocks;
    u.assign((size_t)(cap + 1) * pp.neq * 400, 0.0); up = u; pr = u;
    lp.assign((size_t)(cap + 1) * 8, 0); leaf.assign(cap + 1, 0); ic.assign((size_t)(cap + 1) * 2, 0);
    b.assign((size_t)(cap + 1) * 12, 0);
    m = wxo_mesh{cap, 0, 0, 0, u.data(), up.data(), pr.data(), lp.data(), leaf.data(), ic.data(), b.data()};
    wxo_basegrid((const wxo_params *)&p, &m);
  }
};

static int run_ensemble(int rounds, unsigned seed, int groups) {
  const int n = 3, neq = 4;
  const int nrx[n] = {8, 4, 2}, nry[n] = {8, 2, 6}, nl[n] = {5, 3, 4}, caps[n] = {30000, 4000, 10000};
  std::vector<wx_params> P(n);
  std::vector<int> base(n), grp(n);
  if (groups < 1 || groups > n) { fprintf(stderr, "groups %d\n", groups); return 2; }
  for (int k = 0, i = 0; k < groups; ++k)
    for (int c = 0; c < n / groups + (k < n % groups ? 1 : 0); ++c) grp[i++] = k;
  std::vector<int> gslots(groups, 0);
  for (int i = 0; i < n; ++i) {
    wx_params &p = P[i];
    p = wx_params{};
    p.neq = neq; p.nlevs = nl[i]; p.nbrootx = nrx[i]; p.nbrooty = nry[i]; p.nblocks = caps[i];
    p.cv = 1.5; p.gamma = 5. / 3.; p.cfl = 0.4; p.eta = 1e-3; p.tol = 0.05; p.tol2 = 1.0;
    p.compat = 3; p.Tempsc = 1; p.Psc = 1; p.tsc = 1;
    for (int l = 0; l < nl[i]; ++l) { p.dx[l] = 1.0 / (16 * nrx[i]) / (1 << l); p.dy[l] = 1.0 / (16 * nry[i]) / (1 << l); }
    p.jet[11] = grp[i];
    base[i] = gslots[grp[i]];
    gslots[grp[i]] += caps[i] + 1;
  }
  wx_ens *e = nullptr;
  if (wx_ens_create_groups(P.data(), n, groups, &e)) { fprintf(stderr, "ens create failed\n"); return 2; }
  std::vector<OMesh> om(n);
  for (int i = 0; i < n; ++i) {
    std::vector<double> u((size_t)(nrx[i] * nry[i] + 1) * neq * 400, 1.0);
    if (wx_ens_basegrid(e, i, u.data(), nrx[i] * nry[i])) { fprintf(stderr, "ens basegrid failed\n"); return 2; }
    om[i].init(P[i]);
  }
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> U01(0.0, 1.0);
  int total_ops = 0;
  for (int r = 0; r < rounds; ++r) {
    std::vector<std::vector<signed char>> f(groups);
    for (int k = 0; k < groups; ++k) f[k].assign(gslots[k], 0);
    std::vector<std::vector<int32_t>> up(n), dn(n);
    for (int i = 0; i < n; ++i) {
      wxo_mesh &m = om[i].m;
      up[i].assign(caps[i] + 1, 0);
      dn[i].assign(caps[i] + 1, 0);
      const double pup = ((r + 7 * i) % 40 < 20) ? 0.06 : 0.01, pdn = ((r + 7 * i) % 40 < 20) ? 0.3 : 0.8;
      for (int nb = 1; nb <= m.nbmax; ++nb) {
        if (!om[i].leaf[nb]) continue;
        const int lev = om[i].lp[(size_t)nb * 8];
        const double x = U01(rng);
        if (x < pup && lev < nl[i]) { f[grp[i]][base[i] + nb] = 1; up[i][nb] = 1; }
        else if (x < pup + pdn) { f[grp[i]][base[i] + nb] = 2; dn[i][nb] = 1; }
      }
    }
    for (int k = 0; k < groups; ++k) wxsim_set_group_flags(k, f[k].data(), gslots[k]);
    double z = 0;
    if (wx_ens_step(e, 1, 1, &z)) {
      char b[512];
      wx_last_error(b, 512);
      fprintf(stderr, "round %d: wx_ens_step failed: %s\n", r, b);
      return 3;
    }
    std::vector<int> ops(6 * 60000);
    total_ops += wxsim_ops(ops.data(), 60000);
    for (int i = 0; i < n; ++i) {
      wxo_mesh &m = om[i].m;
      int oerr = wxo_admesh_flags((const wxo_params *)&P[i], &m, up[i].data(), dn[i].data());
      if (oerr) { fprintf(stderr, "round %d: oracle admesh error %d\n", r, oerr); return 3; }
      const int ld = m.nbmax;
      std::vector<int32_t> flp((size_t)8 * ld), fleaf(ld), fic((size_t)2 * ld);
      int32_t nbmax = 0, nbleafs = 0;
      if (wx_ens_get_tree(e, i, flp.data(), fleaf.data(), fic.data(), ld, &nbmax, &nbleafs)) { fprintf(stderr, "get_tree\n"); return 3; }
      if (nbmax != m.nbmax || nbleafs != m.nbleafs) {
        fprintf(stderr, "round %d instance %d: nbmax %d/%d nbleafs %d/%d\n", r, i, nbmax, m.nbmax, nbleafs, m.nbleafs);
        return 1;
      }
      for (int nb = 1; nb <= nbmax; ++nb) {
        for (int k = 1; k <= 8; ++k)
          if (flp[(size_t)(k - 1) * ld + nb - 1] != om[i].lp[(size_t)nb * 8 + k - 1]) {
            fprintf(stderr, "round %d instance %d: lp(%d,%d) differs\n", r, i, nb, k);
            return 1;
          }
        if (fleaf[nb - 1] != om[i].leaf[nb] || fic[nb - 1] != om[i].ic[(size_t)nb * 2] ||
            fic[(size_t)ld + nb - 1] != om[i].ic[(size_t)nb * 2 + 1]) {
          fprintf(stderr, "round %d instance %d: leaf/icoord of %d differ\n", r, i, nb);
          return 1;
        }
      }
    }
  }
  printf("hostsim ensemble: %d instances in %d groups, %d rounds, %d mesh ops -- trees identical\n", n, groups, rounds,
         total_ops);
  wx_ens_destroy(e);
  return 0;
}
