// COCR_FP — drop-in replacement of the reference's COCR_FP process on the GPU.
//
// Process contract (Code/COCR_FP/HarmVect3d.cpp:118-365, launched by Code/CalcStarter/
// main.cpp:302 with cwd = Calculation/Nu{f}_{X|Y} and no arguments):
//   reads  sig3d|Sig3d, mu3d|mu3D, dpr3d|dpr3D, inftry.dat, nver.dat, nvkat.dat, xyz.dat,
//          l13d.dat|L13d.dat (vec_prep_data.cpp:55-187), tsize3d_.dat, nodesforedges.dat,
//          edges.dat (HarmVect3d.cpp:36-112), kuslau, pr, ig, idi, di, jg, ijg, gg
//          (HarmVect3d.cpp:181-229), tsize3d.dat (+ ig3d/jg3d/gg3d.dat when non-zero),
//          nthreads.txt (accepted, unused on the GPU)
//   writes v3.dat (2*nb float64, re/im per edge), kit ("r=%e\teps=%e\tkit=%d\ttime=%lf sec"),
//          logharm3dCalc
//   exit 0 even when the solver does not converge (HarmVect3d.cpp:322,364); 1 on input errors.
//
// Options (none are needed for the drop-in use):
//   --dir D       work in directory D instead of the current directory
//   --mixed       FP32 inner COCR with FP64 refinement (stop on the true FP64 residual)
//   --strict      exit 2 when the solver returns a non-zero request code
//   --batch D...  solve the systems of several directories (one mesh, several frequencies /
//                 polarizations) in one batched launch; K and S are recovered from the
//                 directories' matrices and their "nu" files
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mtcocr.h"

namespace {

constexpr double kPi = 3.1415926535897932384626433832795;

struct Inputs {
  int kuzlov = 0, kpar = 0, kt1 = 0, n_materials = 0;
  std::vector<double> sigma3d;
  std::vector<int> nver, nvkat, l13d, edges;
  int n_edges = 0, n_dc = 0, n_c = 0;
  int n = 0, maxiter = 0;
  double eps = 0;
  std::vector<int> ig, idi, jg, ijg;
  std::vector<double> pr, di, gg;
  int tsize3d = 0;
  std::vector<int> ig_t, jg_t;
  std::vector<double> gg_t;
  std::vector<int> is_node_bound;
  double nu = 0;  // only for --batch
};

std::string join(const std::string& dir, const std::string& name) {
  return dir.empty() || dir == "." ? name : dir + "/" + name;
}

// The reference opens these by their lower-case names on a case-insensitive file system while
// RegularMeshBuilder writes capitalised ones (RegularMeshBuilder/main.cpp:2124,2146-2162).
FILE* open_any(const std::string& dir, std::initializer_list<const char*> names,
               const char* mode, std::string* used = nullptr) {
  for (const char* nm : names) {
    FILE* f = std::fopen(join(dir, nm).c_str(), mode);
    if (f) {
      if (used) *used = nm;
      return f;
    }
  }
  return nullptr;
}

bool fail(std::ofstream& log, const std::string& msg) {
  log << msg << "\n" << std::flush;
  std::fprintf(stderr, "COCR_FP: %s\n", msg.c_str());
  return false;
}

template <class T>
bool read_bin(const std::string& dir, std::initializer_list<const char*> names, T* dst,
              size_t count, std::ofstream& log) {
  std::string used;
  FILE* f = open_any(dir, names, "rb", &used);
  if (!f) return fail(log, std::string("Cannot open file ") + *names.begin());
  const size_t got = std::fread(dst, sizeof(T), count, f);
  std::fclose(f);
  if (got != count) return fail(log, "Cannot read binary file " + used);
  return true;
}

bool read_inputs(const std::string& dir, Inputs& in, std::ofstream& log, bool need_nu) {
  FILE* f;
  // materials: the number of materials is the largest id in sig3d (vec_prep_data.cpp:70-81)
  f = open_any(dir, {"sig3d", "Sig3d", "SIG3D"}, "r");
  if (!f) return fail(log, "CANNOT OPEN FILE \"sig3d\"");
  {
    int id;
    double a, b;
    std::vector<std::pair<int, double>> rows;
    while (std::fscanf(f, "%d %lf %lf", &id, &a, &b) == 3) {
      rows.push_back({id, a});
      if (id > in.n_materials) in.n_materials = id;
    }
    std::fclose(f);
    in.sigma3d.assign(in.n_materials, 0.0);
    for (int i = 0; i < (int)rows.size() && i < in.n_materials; ++i)
      if (rows[i].first >= 1) in.sigma3d[rows[i].first - 1] = rows[i].second;
  }
  if (!(f = open_any(dir, {"mu3d", "mu3D", "Mu3d", "MU3D"}, "r")))
    return fail(log, "CANNOT OPEN FILE \"mu3d\"");
  std::fclose(f);
  if (!(f = open_any(dir, {"dpr3d", "dpr3D", "Dpr3d", "DPR3D"}, "r")))
    return fail(log, "CANNOT OPEN FILE \"dpr3d\"");
  std::fclose(f);
  // inftry.dat: KUZLOV= KPAR= KT1= (in_out.cpp:492-526)
  if (!(f = open_any(dir, {"inftry.dat"}, "r"))) return fail(log, "Cannot open inftry.dat");
  {
    char buf[256];
    while (std::fscanf(f, "%255s", buf) == 1) {
      if (!std::strcmp(buf, "KUZLOV=")) { if (std::fscanf(f, "%d", &in.kuzlov) != 1) break; }
      else if (!std::strcmp(buf, "KPAR=")) { if (std::fscanf(f, "%d", &in.kpar) != 1) break; }
      else if (!std::strcmp(buf, "KT1=")) { if (std::fscanf(f, "%d", &in.kt1) != 1) break; }
    }
    std::fclose(f);
  }
  if (in.kuzlov <= 0 || in.kpar <= 0 || in.kt1 < 0) return fail(log, "bad inftry.dat");
  in.nver.resize((size_t)in.kpar * 14);
  in.nvkat.resize(in.kpar);
  in.l13d.resize(in.kt1);
  if (!read_bin(dir, {"nver.dat"}, in.nver.data(), in.nver.size(), log)) return false;
  if (!read_bin(dir, {"nvkat.dat"}, in.nvkat.data(), in.nvkat.size(), log)) return false;
  if (in.kt1 && !read_bin(dir, {"l13d.dat", "L13d.dat", "L13D.dat"}, in.l13d.data(), in.l13d.size(), log))
    return false;
  {
    std::vector<double> xyz((size_t)in.kuzlov * 3);
    if (!read_bin(dir, {"xyz.dat"}, xyz.data(), xyz.size(), log)) return false;
  }
  for (auto& v : in.nver) v--;
  for (auto& v : in.nvkat) v--;
  for (auto& v : in.l13d) v--;
  // inputvec: tsize3d_.dat (n_dc, n), nodesforedges.dat, edges.dat (HarmVect3d.cpp:36-112)
  if (!(f = open_any(dir, {"tsize3d_.dat"}, "r"))) return fail(log, "Error open file tsize3d_.dat");
  if (std::fscanf(f, "%d %d", &in.n_dc, &in.n_edges) != 2) {
    std::fclose(f);
    return fail(log, "bad tsize3d_.dat");
  }
  std::fclose(f);
  in.n_c = in.n_edges - in.n_dc;
  in.edges.resize((size_t)in.n_edges * 2);
  if (!read_bin(dir, {"nodesforedges.dat"}, in.edges.data(), in.edges.size(), log)) return false;
  for (auto& v : in.edges) v--;
  {
    std::vector<int> ed((size_t)in.kpar * 25);
    if (!read_bin(dir, {"edges.dat"}, ed.data(), ed.size(), log)) return false;
  }
  // kuslau + the SLAE (HarmVect3d.cpp:181-229)
  if (!(f = open_any(dir, {"kuslau"}, "r"))) return fail(log, "Error in open file kuslau");
  if (std::fscanf(f, "%d %lf %d", &in.n, &in.eps, &in.maxiter) != 3) {
    std::fclose(f);
    return fail(log, "bad kuslau");
  }
  std::fclose(f);
  if (in.n != 2 * in.n_c) return fail(log, "kuslau n does not match tsize3d_.dat");
  const int nb = in.n_c;
  in.pr.resize(in.n);
  in.ig.resize(nb + 1);
  in.idi.resize(nb + 1);
  if (!read_bin(dir, {"pr"}, in.pr.data(), in.pr.size(), log)) return false;
  if (!read_bin(dir, {"ig"}, in.ig.data(), in.ig.size(), log)) return false;
  if (!read_bin(dir, {"idi"}, in.idi.data(), in.idi.size(), log)) return false;
  for (auto& v : in.ig) v--;
  for (auto& v : in.idi) v--;
  const int nnz = in.ig[nb];
  if (nnz < 0 || in.idi[nb] < 0) return fail(log, "bad ig/idi");
  in.jg.resize(nnz);
  in.ijg.resize(nnz + 1);
  in.di.resize(in.idi[nb]);
  if (!read_bin(dir, {"di"}, in.di.data(), in.di.size(), log)) return false;
  if (!read_bin(dir, {"jg"}, in.jg.data(), in.jg.size(), log)) return false;
  if (!read_bin(dir, {"ijg"}, in.ijg.data(), in.ijg.size(), log)) return false;
  for (auto& v : in.jg) v--;
  for (auto& v : in.ijg) v--;
  in.gg.resize(in.ijg[nnz]);
  if (!read_bin(dir, {"gg"}, in.gg.data(), in.gg.size(), log)) return false;
  // optional hanging-node T-matrix (HarmVect3d.cpp:255-301)
  in.tsize3d = 0;
  if ((f = open_any(dir, {"tsize3d.dat"}, "r"))) {
    if (std::fscanf(f, "%d", &in.tsize3d) != 1) in.tsize3d = 0;
    std::fclose(f);
  }
  const int n_nodes_c = in.kuzlov - in.tsize3d;
  if (in.tsize3d > 0) {
    in.ig_t.assign(in.kuzlov + 1, 0);
    if (!read_bin(dir, {"ig3d.dat"}, in.ig_t.data() + n_nodes_c, in.tsize3d + 1, log)) return false;
    for (int i = 0; i < n_nodes_c; ++i) in.ig_t[i] = 1;
    for (auto& v : in.ig_t) v--;
    const int nt = in.ig_t[in.kuzlov];
    in.jg_t.resize(nt);
    in.gg_t.resize(nt);
    if (!read_bin(dir, {"gg3d.dat"}, in.gg_t.data(), nt, log)) return false;
    if (!read_bin(dir, {"jg3d.dat"}, in.jg_t.data(), nt, log)) return false;
    for (auto& v : in.jg_t) v--;
  }
  in.is_node_bound.assign(in.kuzlov, 0);
  for (int v : in.l13d)
    if (v >= 0 && v < in.kuzlov) in.is_node_bound[v] = 1;
  if (need_nu) {
    if (!(f = open_any(dir, {"nu"}, "r"))) return fail(log, "Cannot open nu");
    if (std::fscanf(f, "%lf", &in.nu) != 1) {
      std::fclose(f);
      return fail(log, "bad nu");
    }
    std::fclose(f);
  }
  return true;
}

// kit (WriteKitChrono, base_solver.cpp:181-200) and v3.dat (HarmVect3d.cpp:336); `seconds` is
// the solve time written to kit
bool write_outputs(const std::string& dir, const Inputs& in, const std::vector<double>& x,
                   const mtcocr_stats& st, double seconds) {
  FILE* f = std::fopen(join(dir, "kit").c_str(), "w");
  if (f) {
    std::fprintf(f, "r=%e\teps=%e\tkit=%d\t", st.residual_rel, in.eps, st.iter);
    std::fprintf(f, "time=%lf sec\n", seconds);
    std::fclose(f);
  }
  f = std::fopen(join(dir, "v3.dat").c_str(), "wb");
  if (!f) return false;
  const size_t w = std::fwrite(x.data(), sizeof(double), x.size(), f);
  std::fclose(f);
  return w == x.size();
}

int set_gradient(mtcocr_handle* h, Inputs& in) {
  std::vector<double> sig(in.n_materials);
  for (int i = 0; i < in.n_materials; ++i) sig[i] = (in.sigma3d[i] == 0) ? 0 : 1;
  const int n_nodes_c = in.kuzlov - in.tsize3d;
  return mtcocr_set_gradient_mesh(h, in.kpar, in.n_c, n_nodes_c, in.nvkat.data(), in.nver.data(),
                                  sig.data(), in.edges.data(),
                                  in.tsize3d ? in.ig_t.data() : nullptr,
                                  in.tsize3d ? in.jg_t.data() : nullptr,
                                  in.tsize3d ? in.gg_t.data() : nullptr, in.is_node_bound.data());
}

// the device's locality order of the mesh: the grid order of a box mesh (which also enables
// the stencil SpMV for batches), else the pencil order of a structured hexahedral one, else the
// order given (hanging nodes)
void set_order(mtcocr_handle* h, Inputs& in) {
  if (in.tsize3d) return;
  if (mtcocr_set_mesh_order(h, in.kpar, in.kuzlov, in.nver.data(), in.n_c, in.edges.data(),
                            MTCOCR_ORDER_GRID) != MTCOCR_OK)
    (void)mtcocr_set_mesh_order(h, in.kpar, in.kuzlov, in.nver.data(), in.n_c, in.edges.data(),
                                16);
}

int run_single(const std::string& dir, bool mixed, bool strict) {
  auto t0 = std::chrono::steady_clock::now();
  std::ofstream log(join(dir, "logharm3dCalc"));
  auto stamp = [&](const char* what) {
    log << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << " : "
        << what << std::endl;
  };
  stamp("Reading SLAE");
  Inputs in;
  if (!read_inputs(dir, in, log, false)) return 1;
  int dev = 0;
  if (const char* e = std::getenv("MTCOCR_DEVICE")) dev = std::atoi(e);
  mtcocr_handle* h = nullptr;
  int rc = mtcocr_create(dev, &h);
  if (rc) {
    log << "device error: " << mtcocr_error_string(rc) << std::endl;
    std::fprintf(stderr, "COCR_FP: %s\n", mtcocr_error_string(rc));
    return 1;
  }
  set_order(h, in);
  rc = mtcocr_set_matrix_block2x2(h, in.n_c, in.ig.data(), in.jg.data(), in.idi.data(),
                                  in.ijg.data(), in.di.data(), in.gg.data());
  if (!rc) rc = set_gradient(h, in);
  if (rc) {
    log << "setup error: " << mtcocr_error_string(rc) << std::endl;
    std::fprintf(stderr, "COCR_FP: %s\n", mtcocr_error_string(rc));
    mtcocr_destroy(h);
    return 1;
  }
  stamp("Solving SLAE");
  std::vector<double> u(in.n, 0.0);  // x0 = 0 (HarmVect3d.cpp:235)
  mtcocr_stats st;
  std::memset(&st, 0, sizeof(st));
  rc = mtcocr_solve(h, in.pr.data(), u.data(), in.eps, in.maxiter,
                    mixed ? MTCOCR_PREC_MIXED : MTCOCR_PREC_FP64, &st);
  mtcocr_destroy(h);
  if (rc <= MTCOCR_ERR_ARG && rc != MTCOCR_ERR_NOTCONV) {
    log << "solver error: " << mtcocr_error_string(rc) << std::endl;
    std::fprintf(stderr, "COCR_FP: %s\n", mtcocr_error_string(rc));
    return 1;
  }
  log << st.iter << "    " << std::scientific << st.residual_rel << std::endl;
  stamp("Writing result");
  if (!write_outputs(dir, in, u, st, st.seconds)) return 1;
  stamp("All done");
  std::printf("time=%lf\n", st.seconds);
  if (strict && rc != 0) return 2;
  return 0;
}

// Several directories of one mesh: A_d = K + i*omega_d*S with K = Re(A_d), S = Im(A_d)/omega_d.
int run_batch(const std::vector<std::string>& dirs, bool mixed, bool strict) {
  const int nsys = (int)dirs.size();
  if (nsys < 1 || nsys > 64) {
    std::fprintf(stderr, "COCR_FP: --batch takes 1..64 directories\n");
    return 1;
  }
  std::vector<Inputs> ins(nsys);
  std::vector<std::ofstream> logs;
  for (int s = 0; s < nsys; ++s) {
    logs.emplace_back(join(dirs[s], "logharm3dCalc"));
    if (!read_inputs(dirs[s], ins[s], logs[s], true)) return 1;
  }
  const Inputs& a = ins[0];
  const int nb = a.n_c;
  for (int s = 1; s < nsys; ++s) {
    if (ins[s].n_c != nb || ins[s].ig != a.ig || ins[s].jg != a.jg || ins[s].ijg != a.ijg ||
        ins[s].idi != a.idi) {
      std::fprintf(stderr, "COCR_FP: --batch directories do not share one matrix portrait\n");
      return 1;
    }
    // one batch runs with one tolerance and iteration limit (kuslau)
    if (ins[s].eps != a.eps || ins[s].maxiter != a.maxiter) {
      std::fprintf(stderr, "COCR_FP: --batch: %s has eps/maxiter %g/%d, %s has %g/%d\n",
                   dirs[s].c_str(), ins[s].eps, ins[s].maxiter, dirs[0].c_str(), a.eps,
                   a.maxiter);
      return 1;
    }
  }
  // split the first directory's matrix into K and S; verify every other one against it
  const double w0 = 2.0 * kPi * a.nu;
  const int nnz = a.ig[nb];
  std::vector<double> kd(nb), sd(nb), kg(nnz), sg(nnz);
  for (int i = 0; i < nb; ++i) {
    const int sz = a.idi[i + 1] - a.idi[i];
    kd[i] = a.di[a.idi[i]];
    sd[i] = sz == 2 ? a.di[a.idi[i] + 1] / w0 : 0.0;
  }
  for (int k = 0; k < nnz; ++k) {
    const int sz = a.ijg[k + 1] - a.ijg[k];
    kg[k] = a.gg[a.ijg[k]];
    sg[k] = sz == 2 ? a.gg[a.ijg[k] + 1] / w0 : 0.0;
  }
  // every block of every directory (diagonal and off-diagonal) must be K + i*omega_s*S
  auto matches = [](double re, double im, double k, double s_om) {
    return std::fabs(re - k) <= 1e-12 * std::fabs(k) + 1e-300 &&
           std::fabs(im - s_om) <= 1e-10 * std::fabs(im) + 1e-300;
  };
  std::vector<double> omega(nsys);
  for (int s = 0; s < nsys; ++s) {
    omega[s] = 2.0 * kPi * ins[s].nu;
    const Inputs& b = ins[s];
    bool ok = true;
    for (int i = 0; ok && i < nb; ++i) {
      const int sz = b.idi[i + 1] - b.idi[i];
      ok = matches(b.di[b.idi[i]], sz == 2 ? b.di[b.idi[i] + 1] : 0.0, kd[i], omega[s] * sd[i]);
    }
    for (int k = 0; ok && k < nnz; ++k) {
      const int sz = b.ijg[k + 1] - b.ijg[k];
      ok = matches(b.gg[b.ijg[k]], sz == 2 ? b.gg[b.ijg[k] + 1] : 0.0, kg[k], omega[s] * sg[k]);
    }
    if (!ok) {
      std::fprintf(stderr, "COCR_FP: --batch: %s is not K + i*omega*S of %s\n", dirs[s].c_str(),
                   dirs[0].c_str());
      return 1;
    }
  }
  int dev = 0;
  if (const char* e = std::getenv("MTCOCR_DEVICE")) dev = std::atoi(e);
  mtcocr_handle* h = nullptr;
  int rc = mtcocr_create(dev, &h);
  if (!rc) set_order(h, ins[0]);
  if (!rc) rc = mtcocr_set_matrix_ks(h, nb, a.ig.data(), a.jg.data(), kd.data(), kg.data(),
                                     sd.data(), sg.data());
  if (!rc) rc = set_gradient(h, ins[0]);
  if (rc) {
    std::fprintf(stderr, "COCR_FP: %s\n", mtcocr_error_string(rc));
    if (h) mtcocr_destroy(h);
    return 1;
  }
  std::vector<double> pr((size_t)nsys * 2 * nb), x((size_t)nsys * 2 * nb, 0.0);
  for (int s = 0; s < nsys; ++s) std::memcpy(&pr[(size_t)s * 2 * nb], ins[s].pr.data(), sizeof(double) * 2 * nb);
  std::vector<mtcocr_stats> st(nsys);
  rc = mtcocr_solve_batch(h, nsys, omega.data(), pr.data(), x.data(), a.eps, a.maxiter,
                          mixed ? MTCOCR_PREC_MIXED : MTCOCR_PREC_FP64, st.data());
  mtcocr_destroy(h);
  if (rc <= MTCOCR_ERR_ARG && rc != MTCOCR_ERR_NOTCONV) {
    std::fprintf(stderr, "COCR_FP: %s\n", mtcocr_error_string(rc));
    return 1;
  }
  // The systems of a batch share every iteration's launches, so a system has no time of its
  // own: each kit gets the batch's device time divided by the number of systems (its share),
  // and logharm3dCalc the whole batch's time.
  const double batch_s = nsys > 0 ? st[0].seconds : 0.0;
  for (int s = 0; s < nsys; ++s) {
    std::vector<double> u(x.begin() + (size_t)s * 2 * nb, x.begin() + (size_t)(s + 1) * 2 * nb);
    logs[s] << st[s].iter << "    " << std::scientific << st[s].residual_rel << std::endl;
    logs[s] << "batch of " << nsys << " systems: " << std::fixed << batch_s << " sec" << std::endl;
    if (!write_outputs(dirs[s], ins[s], u, st[s], batch_s / nsys)) return 1;
  }
  std::printf("batch time=%lf sec for %d systems\n", batch_s, nsys);
  if (strict && rc != 0) return 2;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::string dir = ".";
  bool mixed = false, strict = false;
  std::vector<std::string> batch;
  bool in_batch = false;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--dir" && i + 1 < argc) {
      dir = argv[++i];
      in_batch = false;
    } else if (a == "--mixed") {
      mixed = true;
    } else if (a == "--strict") {
      strict = true;
    } else if (a == "--batch") {
      in_batch = true;
    } else if (in_batch) {
      batch.push_back(a);
    } else {
      std::fprintf(stderr, "usage: COCR_FP [--dir D] [--mixed] [--strict] [--batch D1 D2 ...]\n");
      return 1;
    }
  }
  if (!batch.empty()) return run_batch(batch, mixed, strict);
  return run_single(dir, mixed, strict);
}
