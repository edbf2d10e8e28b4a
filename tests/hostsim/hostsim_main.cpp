// Drives the product's host runtime (wx_host.cpp built against the host-only
// HIP stand-in) through many admesh rounds with seeded refinement flags, and
// checks its tree (lp, leaf, icoord, nbmax, nbleafs) against the oracle's
// admesh given the same flags after every round.  Built with ASan/UBSan.
// Usage: hostsim_main <rounds> <seed> [ens].  With "ens", three instances of
// different root layouts, depths and capacities run as one wx_ens ensemble
// and each instance's tree is compared with its own oracle mesh.  The
// single-mesh run cycles the table builder through its modes (host list,
// device builder with either scan, choice by size) and checks the tables it
// holds after every round against a host compilation of the same tree.
// hostsim_main 0 0 writebin STATE DIR uploads the state in the file STATE
// and writes its BIN pair into DIR with wx_write_bin (write_bin_of_state).
// TEST SCAFFOLDING.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include <cstring>

#include "../../include/wx_capi.h"
#include "../../oracle/wxoracle.h"
#include "../../walicxe2d_amd/csrc/wx_runtime.h"
#include "hip/hip_runtime.h"

extern "C" void wxsim_set_flags(const signed char *f, int n);
extern "C" int wxsim_ops(int *out, int cap);
extern "C" void wxsim_set_group_flags(int group, const signed char *f, int n);

static int run_ensemble(int rounds, unsigned seed, int groups);
static int bin_roundtrip(wx_ctx *c, const wx_params &p, const std::string &dir);
static int write_bin_of_state(const char *state, const char *dir);

// The context's current tables (wx_get_tables) against compile_leaves +
// locatebounds + compile_ops on a copy of its tree.
static int check_tables(wx_ctx *c, const wx_params &p, int r) {
  int32_t n = 0, nbmax = 0, nbleafs = 0;
  if (wx_get_tables(c, &n, nullptr, nullptr, nullptr, 0)) return 3;
  std::vector<int32_t> rec((size_t)n * 8), o2((size_t)n * 16), o1((size_t)n * 16);
  if (wx_get_tables(c, &n, rec.data(), o2.data(), o1.data(), n)) return 3;
  if (wx_info(c, &nbmax, nullptr, nullptr)) return 3;
  const int ld = nbmax > 0 ? nbmax : 1;
  std::vector<int32_t> lp((size_t)8 * ld), leaf(ld), ic((size_t)2 * ld);
  if (wx_get_tree(c, lp.data(), leaf.data(), ic.data(), ld, &nbmax, &nbleafs)) return 3;
  wx::Mesh m;
  m.init(p);
  if (wx::set_tree(&m, lp.data(), leaf.data(), ic.data(), ld, nbmax, nbleafs, nullptr) || wx::refresh_tables(&m))
    return 3;
  if ((size_t)n != m.leaves.size()) { fprintf(stderr, "round %d: %d leaves vs %zu\n", r, n, m.leaves.size()); return 1; }
  if (memcmp(rec.data(), m.leaves.data(), (size_t)n * 32) || memcmp(o2.data(), m.opsII.data(), (size_t)n * 64) ||
      memcmp(o1.data(), m.opsI.data(), (size_t)n * 64)) {
    fprintf(stderr, "round %d: tables differ from the host compilation\n", r);
    return 1;
  }
  return 0;
}

int main(int argc, char **argv) {
  const int rounds = argc > 1 ? atoi(argv[1]) : 200;
  const unsigned seed = argc > 2 ? (unsigned)atoi(argv[2]) : 1u;
  if (argc > 3 && std::string(argv[3]) == "ens") return run_ensemble(rounds, seed, argc > 4 ? atoi(argv[4]) : 1);
  if (argc > 5 && std::string(argv[3]) == "writebin") return write_bin_of_state(argv[4], argv[5]);
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
    if (wx_set_builder(c, (r / 7) % 4)) { fprintf(stderr, "set_builder failed\n"); return 3; }
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
    if (int rc = check_tables(c, p, r)) return rc;
    int32_t nint = 0;
    if (wx_locatebounds(c, &nint)) { fprintf(stderr, "locatebounds failed\n"); return 3; }
    if (nint != wxo_locatebounds((const wxo_params *)&p, &m)) { fprintf(stderr, "round %d: list length differs\n", r); return 1; }
  }
  printf("hostsim: %d rounds, %d mesh ops, final nbmax %d leaves %d -- trees identical, tables identical\n", rounds,
         total_ops, m.nbmax, m.nbleafs);
  if (argc > 4 && std::string(argv[3]) == "bin") {
    const int rc = bin_roundtrip(c, p, argv[4]);
    if (rc) return rc;
  }
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
    // seeds 1 mod 4: reflective bottoms, so no instance is in the OUTFYB mode
    // and fill I's ops are fill II's for the whole ensemble
    if (seed % 4 == 1) p.bc[2] = 1;
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
  // steps replayed from captured graphs (re-captured as the queues grow and
  // shrink); a third of the way the fill switches to a persistent grid and
  // back at two thirds, which drops and re-captures the graphs again
  if (wx_ens_set_graph(e, 1)) { fprintf(stderr, "set_graph failed\n"); return 2; }
  for (int r = 0; r < rounds; ++r) {
    if (r == rounds / 3 && wx_ens_set_launch(e, 2, 64, 256)) { fprintf(stderr, "set_launch failed\n"); return 2; }
    if (r == 2 * rounds / 3 && wx_ens_set_launch(e, 2, 0, 256)) { fprintf(stderr, "set_launch failed\n"); return 2; }
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
  printf("hostsim ensemble: %d instances in %d groups, %d rounds, %d mesh ops, %d graph captures, %d replays -- "
         "trees identical\n", n, groups, rounds, total_ops, hipsim::captures, hipsim::replays);
  wx_ens_destroy(e);
  return 0;
}

// ---------------------------------------------------------------------
// BIN checkpoints: the final mesh with a tile pattern (every slot, ghosts
// included) is written by wx_write_bin, read back into a fresh context by
// wx_read_bin and compared; the expected tree and tiles go to raw files
// (expect_tree.i32: nbmax, nbleafs, lp(8 x nbmax), leaf, icoord(2 x
// nbmax); expect_u.f64: (nbmax+1) tiles [var][j][i]) for the Python reader.
static int bin_roundtrip(wx_ctx *c, const wx_params &p, const std::string &dir) {
  int32_t nbmax = 0, nbleafs = 0;
  wx_info(c, &nbmax, &nbleafs, nullptr);
  const int ld = nbmax, neq = p.neq;
  std::vector<int32_t> lp((size_t)8 * ld), leaf(ld), ic((size_t)2 * ld);
  int32_t a = 0, b = 0;
  if (wx_get_tree(c, lp.data(), leaf.data(), ic.data(), ld, &a, &b)) return 3;
  std::vector<double> u((size_t)(nbmax + 1) * neq * 400);
  for (size_t n = 0; n < u.size(); ++n) u[n] = 1.0 + 1e-3 * (double)(n % 100003) + (double)(n / 400) * 1e-9;
  if (wx_upload(c, u.data(), WX_LAYOUT_BLOCK, lp.data(), leaf.data(), ic.data(), ld, nbmax, nbleafs)) return 3;
  const std::string pp = dir + "/pointers000.001.bin", bp = dir + "/blocks000.001.bin";
  if (wx_write_bin(c, pp.c_str(), bp.c_str(), -1)) {
    char e[512];
    wx_last_error(e, 512);
    fprintf(stderr, "write_bin: %s\n", e);
    return 3;
  }
  FILE *f = fopen((dir + "/expect_tree.i32").c_str(), "wb");
  const int32_t head[2] = {nbmax, nbleafs};
  fwrite(head, 4, 2, f);
  fwrite(lp.data(), 4, lp.size(), f);
  fwrite(leaf.data(), 4, leaf.size(), f);
  fwrite(ic.data(), 4, ic.size(), f);
  fclose(f);
  f = fopen((dir + "/expect_u.f64").c_str(), "wb");
  fwrite(u.data(), 8, u.size(), f);
  fclose(f);
  wx_ctx *c2 = nullptr;
  if (wx_create(&p, &c2)) return 3;
  const char *pa[1] = {pp.c_str()}, *ba[1] = {bp.c_str()};
  if (wx_read_bin(c2, 1, pa, ba)) {
    char e[512];
    wx_last_error(e, 512);
    fprintf(stderr, "read_bin: %s\n", e);
    return 3;
  }
  std::vector<int32_t> lp2(lp.size()), leaf2(leaf.size()), ic2(ic.size());
  if (wx_get_tree(c2, lp2.data(), leaf2.data(), ic2.data(), ld, &a, &b)) return 3;
  if (a != nbmax || b != nbleafs || lp2 != lp || leaf2 != leaf || ic2 != ic) {
    fprintf(stderr, "bin: tree differs after the round trip\n");
    return 1;
  }
  std::vector<double> u2(u.size());
  if (wx_download(c2, u2.data(), nullptr, WX_LAYOUT_BLOCK, 0)) return 3;
  for (int nb = 1; nb <= nbmax; ++nb) {
    if (!leaf[nb - 1]) continue;
    for (size_t q = 0; q < (size_t)neq * 400; ++q)
      if (u2[(size_t)nb * neq * 400 + q] != u[(size_t)nb * neq * 400 + q]) {
        fprintf(stderr, "bin: tile of leaf %d differs at %zu\n", nb, q);
        return 1;
      }
  }
  wx_destroy(c2);
  printf("bin: %d blocks, %d leaves written and read back identical\n", nbmax, nbleafs);
  return 0;
}

// ---------------------------------------------------------------------
// writebin: STATE holds int32 neq, nblocks, nlevs, nbmax, nbleafs,
// logical_true, then lp (8 per ID, IDs 1..nbmax), leaf (nbmax), icoord (2
// per ID), then float64 tiles 0..nbmax [var][j][i]; the context gets
// capacity nblocks, the state is uploaded and written by wx_write_bin as
// DIR/pointers000.001.bin and DIR/blocks000.001.bin.
static int write_bin_of_state(const char *state, const char *dir) {
  FILE *f = fopen(state, "rb");
  if (!f) { fprintf(stderr, "writebin: cannot open %s\n", state); return 2; }
  int32_t h[6];
  if (fread(h, 4, 6, f) != 6) { fclose(f); return 2; }
  const int neq = h[0], cap = h[1], nlevs = h[2], nbmax = h[3], nbleafs = h[4], ltrue = h[5];
  std::vector<int32_t> lpr((size_t)8 * nbmax), leaf(nbmax), icr((size_t)2 * nbmax);
  std::vector<double> u((size_t)(nbmax + 1) * neq * 400);
  bool ok = fread(lpr.data(), 4, lpr.size(), f) == lpr.size() && fread(leaf.data(), 4, leaf.size(), f) == leaf.size() &&
            fread(icr.data(), 4, icr.size(), f) == icr.size() && fread(u.data(), 8, u.size(), f) == u.size();
  fclose(f);
  if (!ok) { fprintf(stderr, "writebin: short state file\n"); return 2; }
  // the C-ABI takes the Fortran column-major lp(ld,8) / icoord(ld,2)
  std::vector<int32_t> lp((size_t)8 * nbmax), ic((size_t)2 * nbmax);
  for (int nb = 0; nb < nbmax; ++nb) {
    for (int k = 0; k < 8; ++k) lp[(size_t)k * nbmax + nb] = lpr[(size_t)nb * 8 + k];
    for (int d = 0; d < 2; ++d) ic[(size_t)d * nbmax + nb] = icr[(size_t)nb * 2 + d];
  }
  wx_params p{};
  p.neq = neq; p.nlevs = nlevs; p.nbrootx = 1; p.nbrooty = 1; p.nblocks = cap;
  p.cv = 1.5; p.gamma = 5. / 3.; p.cfl = 0.4; p.Tempsc = 1; p.Psc = 1; p.tsc = 1; p.tol = 0.05; p.tol2 = 1.0;
  for (int l = 0; l < nlevs; ++l) { p.dx[l] = 1.0 / (1 << l); p.dy[l] = 1.0 / (1 << l); }
  wx_ctx *c = nullptr;
  if (wx_create(&p, &c)) { fprintf(stderr, "writebin: create failed\n"); return 2; }
  char e[512];
  if (wx_upload(c, u.data(), WX_LAYOUT_BLOCK, lp.data(), leaf.data(), ic.data(), nbmax, nbmax, nbleafs)) {
    wx_last_error(e, 512);
    fprintf(stderr, "writebin: upload: %s\n", e);
    return 3;
  }
  const std::string pp = std::string(dir) + "/pointers000.001.bin", bp = std::string(dir) + "/blocks000.001.bin";
  if (wx_write_bin(c, pp.c_str(), bp.c_str(), ltrue)) {
    wx_last_error(e, 512);
    fprintf(stderr, "writebin: %s\n", e);
    return 3;
  }
  wx_destroy(c);
  printf("writebin: %d blocks, %d leaves written\n", nbmax, nbleafs);
  return 0;
}
