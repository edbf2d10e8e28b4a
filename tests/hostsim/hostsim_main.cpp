// Drives the product's host runtime (wx_host.cpp built against the host-only
// HIP stand-in) through many admesh rounds with seeded refinement flags, and
// checks its tree (lp, leaf, icoord, nbmax, nbleafs) against the oracle's
// admesh given the same flags after every round.  Built with ASan/UBSan.
// Usage: hostsim_main <rounds> <seed>.  TEST SCAFFOLDING.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../include/wx_capi.h"
#include "../../oracle/wxoracle.h"

extern "C" void wxsim_set_flags(const signed char *f, int n);
extern "C" int wxsim_ops(int *out, int cap);

int main(int argc, char **argv) {
  const int rounds = argc > 1 ? atoi(argv[1]) : 200;
  const unsigned seed = argc > 2 ? (unsigned)atoi(argv[2]) : 1u;
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
