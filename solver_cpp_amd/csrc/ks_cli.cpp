// ks_cli.cpp - `Solver`, the command-line drop-in for the reference's
// executable (main.cpp:6-30 + collection::readCMD/checkCMD/readMatrix,
// solver_collection.cpp:269-828): same flags and defaults, same CRS input
// directory layout, same ./output/<NAME>_his.txt / _xvec.txt files and the
// stdout lines `|b-ax|2/|b|2 = `, `loop = `, `time = `, `converge` /
// `not converge`.  All solving runs on one GPU through libksolve.
//
// Added flags (no reference counterpart): --inner-precision 32|64,
// --ref-compat BITS, --device N, --json (one JSON result line), --bin-cache
// (load <dir>/CRS.ksb, the binary copy of the text files, when it is current;
// write it otherwise).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <map>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/ksolve.h"

namespace {

struct Opt {
  std::string longname;
  char shortname;
  bool takes_value;
  std::string value;
  bool set = false;
};

const char* kUpper[] = {"", "CG", "BICG", "CR", "GCR", "GMRES", "KSKIPCG", "KSKIPCR", "KSKIPBICG",
                        "VPCG", "VPBICG", "VPCR", "VPGCR", "VPGMRES"};
const char* kLower[] = {"", "cg", "bicg", "cr", "gcr", "gmres", "kskipcg", "kskipcr", "kskipbicg",
                        "vpcg", "vpbicg", "vpcr", "vpgcr", "vpgmres"};

int name_id(const std::string& s) {
  for (int i = 1; i <= 13; ++i)
    if (s == kUpper[i] || s == kLower[i]) return i;
  return 0;
}

bool is_dir(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

void usage(const std::vector<Opt>& opts) {
  std::cerr << "usage: Solver -S <solver> -m <matrix> [options]\n";
  for (auto& o : opts)
    std::cerr << "  " << (o.shortname ? std::string("-") + o.shortname + ", " : std::string("    ")) << "--"
              << o.longname << (o.takes_value ? " <value>  (default " + o.value + ")" : "") << "\n";
}

// The time breakdown in the layout of times::showTimeOnCPU(total, true)
// followed by times::showTimeOnGPU (times.cpp:44-76), which log scrapers read
// line by line ("Dot", "MV", "H2D time", "Memset", "other", data2csv.sh:
// 17-31).  The values are the library's per-class kernel times: Dot = the
// reductions and Gram-Schmidt projections, MV = the SpMVs and the k-skip
// matrix-powers basis, other = the rest of the solve's wall time (updates,
// finishers, launch gaps).  The host does no arithmetic, so the CPU block
// is zero, and nothing crosses PCIe inside the solve, so the transfer lines
// are zero.  Expects std::fixed on the stream, as the reference leaves it.
void print_breakdown(const ks_result& r) {
  const double total = r.seconds;
  const double dot = 1e-3 * (r.prof_ms[KS_K_REDUCE] + r.prof_ms[KS_K_CGS]);
  const double mv = 1e-3 * (r.prof_ms[KS_K_SPMV64] + r.prof_ms[KS_K_SPMV32] + r.prof_ms[KS_K_BASIS]);
  auto line = [&](const char* label, double v) {
    std::cout << label << std::setprecision(6) << v << ", " << std::setprecision(2)
              << (total > 0 ? v / total * 100 : 0.0) << "%" << std::endl;
  };
  std::cout << "Execution time on CPU" << std::endl;
  line("\tDot process time = ", 0.0);
  line("\tMV process time  = ", 0.0);
  line("\t  cons time    = ", 0.0);
  line("\t  dis time    = ", 0.0);
  std::cout << "Execution time on GPU" << std::endl;
  line("\tDot process time  = ", dot);
  line("\tMV process time   = ", mv);
  line("\t  H2D time       = ", 0.0);
  line("\t  D2H time       = ", 0.0);
  line("\t  Memset time    = ", 0.0);
  line("\tother time       = ", total - dot - mv);
  line("\t  Copy time    = ", 0.0);
  line("\t  cons time    = ", 0.0);
  line("\t  dis time    = ", 0.0);
}

}  // namespace

int main(int argc, char** argv) {
  // flag table: solver_collection.cpp:278-307 (defaults :20-49)
  std::vector<Opt> opts = {
      {"L1_Dir_Name", 'N', true, "Matrix"},   {"CRS_Path", 'p', true, "../"},
      {"CRS_Dir_Name", 'd', true, "CRS"},     {"CRS_Matrix_Name", 'm', true, "NONE"},
      {"MM_Path", 'P', true, "../"},          {"MM_Dir_Name", 'D', true, "MM"},
      {"MM_Matrix_Name", 'M', true, "NONE"},  {"Openmp_Thread", 't', true, "8"},
      {"OuterSolver", 'S', true, ""},         {"OuterMaxLoop", 'L', true, "10000"},
      {"OuterEps", 'E', true, "1e-8"},        {"OuterRestart", 'R', true, "1000"},
      {"OuterKskip", 'K', true, "2"},         {"OuterFix", 'F', true, "2"},
      {"InnerSolver", 's', true, "NO"},       {"InnerMaxLoop", 'l', true, "50"},
      {"InnerEps", 'e', true, "1e-1"},        {"InnerRestart", 'r', true, "1000"},
      {"InnerKskip", 'k', true, "2"},         {"InnerFix", 'f', true, "2"},
      {"verbose", 'v', false, ""},            {"cuda", 'c', false, ""},
      {"pinned", 'x', false, ""},             {"multiGPU", 'g', false, ""},
      {"inner-precision", 0, true, "64"},     {"ref-compat", 0, true, "0"},
      {"device", 0, true, "0"},               {"json", 0, false, ""},
      {"bin-cache", 0, false, ""},
      {"help", 'h', false, ""},
  };
  std::map<std::string, Opt*> bylong;
  std::map<char, Opt*> byshort;
  for (auto& o : opts) {
    bylong[o.longname] = &o;
    if (o.shortname) byshort[o.shortname] = &o;
  }
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    Opt* o = nullptr;
    std::string val;
    bool inline_val = false;
    if (a.rfind("--", 0) == 0) {
      std::string nm = a.substr(2);
      size_t eq = nm.find('=');
      if (eq != std::string::npos) { val = nm.substr(eq + 1); nm = nm.substr(0, eq); inline_val = true; }
      auto it = bylong.find(nm);
      if (it == bylong.end()) { std::cerr << "undefined option: --" << nm << "\n"; usage(opts); return 1; }
      o = it->second;
    } else if (a.size() >= 2 && a[0] == '-') {
      auto it = byshort.find(a[1]);
      if (it == byshort.end()) { std::cerr << "undefined option: " << a << "\n"; usage(opts); return 1; }
      o = it->second;
      if (a.size() > 2) { val = a.substr(2); inline_val = true; }
    } else {
      std::cerr << "unexpected argument: " << a << "\n";
      return 1;
    }
    if (o->takes_value) {
      if (!inline_val) {
        if (i + 1 >= argc) { std::cerr << "option needs value: --" << o->longname << "\n"; return 1; }
        val = argv[++i];
      }
      o->value = val;
    }
    o->set = true;
  }
  if (bylong["help"]->set) { usage(opts); return 0; }
  auto get = [&](const char* nm) { return bylong[nm]->value; };
  if (!bylong["OuterSolver"]->set) { std::cerr << "option needs value: --OuterSolver\n"; usage(opts); return 1; }

  // collection::readCMD checks (solver_collection.cpp:333-381)
  if (get("MM_Matrix_Name") == "NONE" && get("CRS_Matrix_Name") == "NONE") {
    std::cerr << "[X]Must set one input source" << std::endl;
    return 1;
  }
  if (get("MM_Matrix_Name") != "NONE" && get("CRS_Matrix_Name") != "NONE") {
    std::cerr << "[X]Only can set one input source" << std::endl;
    return 1;
  }
  if (get("MM_Matrix_Name") != "NONE") {
    std::cerr << "[X]MatrixMarket input is not supported (the reference reads only CRS text)" << std::endl;
    return 1;
  }
  const int outer = name_id(get("OuterSolver"));
  if (!outer) { std::cerr << "[X]OuterSolverName can not found in List" << std::endl; return 1; }
  const int inner = get("InnerSolver") == "NO" ? 0 : name_id(get("InnerSolver"));
  if (get("InnerSolver") != "NO" && !inner) { std::cerr << "[X]InnerSolverName can not found in List" << std::endl; return 1; }
  if (bylong["multiGPU"]->set) {
    std::cerr << "[X]MultiGPU mode is not supported: this build runs on one GPU" << std::endl;
    return 1;
  }

  ks_params p;
  ks_params_default(&p);
  std::snprintf(p.outer, sizeof(p.outer), "%s", get("OuterSolver").c_str());
  std::snprintf(p.inner, sizeof(p.inner), "%s", get("InnerSolver").c_str());
  p.outer_maxloop = std::atoi(get("OuterMaxLoop").c_str());
  p.outer_eps = std::atof(get("OuterEps").c_str());
  p.outer_restart = std::atoi(get("OuterRestart").c_str());
  p.outer_kskip = std::atoi(get("OuterKskip").c_str());
  p.outer_fix = std::atoi(get("OuterFix").c_str());
  p.inner_maxloop = std::atoi(get("InnerMaxLoop").c_str());
  p.inner_eps = std::atof(get("InnerEps").c_str());
  p.inner_restart = std::atoi(get("InnerRestart").c_str());
  p.inner_kskip = std::atoi(get("InnerKskip").c_str());
  p.inner_fix = std::atoi(get("InnerFix").c_str());
  p.inner_precision = std::atoi(get("inner-precision").c_str());
  p.ref_compat = (uint32_t)std::strtoul(get("ref-compat").c_str(), nullptr, 0);
  p.omp_threads = std::atoi(get("Openmp_Thread").c_str());
  p.verbose = bylong["verbose"]->set;
  p.profile = (1 << KS_K_NCLASS) - 1;  // per-class times for the breakdown

  // collection::checkCMD directory walk (solver_collection.cpp:389-502)
  const std::string full = get("CRS_Path") + "/" + get("L1_Dir_Name") + "/" + get("CRS_Dir_Name") + "/" +
                           get("CRS_Matrix_Name");
  if (!is_dir(full)) { std::cerr << "[X]Can't found CRS files Dir -> " << full << std::endl; return 1; }
  int64_t n = 0, nnz = 0;
  if (ks_crs_read_header(full.c_str(), &n, &nnz) != KS_OK) {
    std::cerr << "[X] Error in CRS files (missing file or the three headers disagree) in " << full << std::endl;
    return 1;
  }
  std::vector<int32_t> ptr(n + 1), col(nnz);
  std::vector<double> val(nnz), b(n), x(n);
  // --bin-cache: <dir>/CRS.ksb is used when its stamp matches the text
  // files, and (re)written from the text otherwise
  bool from_cache = false;
  const bool use_cache = bylong["bin-cache"]->set;
  const std::string cache = full + "/CRS.ksb";
  uint64_t stamp = 0;
  if (use_cache && ks_crs_stamp(full.c_str(), &stamp) == KS_OK) {
    int64_t cn = 0, cnnz = 0;
    uint64_t cst = 0;
    if (ks_crsbin_read_header(cache.c_str(), &cn, &cnnz, &cst) == KS_OK && cn == n && cnnz == nnz &&
        cst == stamp)
      from_cache = ks_crsbin_read(cache.c_str(), ptr.data(), col.data(), val.data(), b.data(), x.data()) == KS_OK;
  }
  if (!from_cache) {
    if (ks_crs_read(full.c_str(), ptr.data(), col.data(), val.data(), b.data(), x.data()) != KS_OK) {
      std::cerr << "[X] Read CRS files fail" << std::endl;
      return 1;
    }
    if (use_cache && stamp &&
        ks_crsbin_write(cache.c_str(), n, nnz, ptr.data(), col.data(), val.data(), b.data(), x.data(), stamp) !=
            KS_OK)
      std::cerr << "[!] could not write the binary cache " << cache << std::endl;
  }
  if (!is_dir("./output")) { std::cerr << "File open error" << std::endl; return 1; }
  // collection::showCMD (solver_collection.cpp:603-651).  Every solve runs on
  // the GPU, so "CUDA : On" whether or not -c was given; one GPU is used.
  const char* rule = "-------------";
  std::cout << "=========================================================" << std::endl;
  std::cout << "InputSource : CRS" << (from_cache ? " (binary cache)" : "") << std::endl
            << "FullPath : " << full << std::endl;
  std::cout << "N : " << n << std::endl << "NNZ : " << nnz << std::endl;
  std::cout << rule << std::endl;
  std::cout << "OpenMP thread : " << p.omp_threads << std::endl;
  std::cout << "GPU Num : 1" << std::endl;
  std::cout << rule << std::endl;
  std::cout << "VP : " << (inner ? "On" : "Off") << std::endl;
  std::cout << "CUDA : On" << std::endl;
  std::cout << "Pinned : Off" << std::endl;
  std::cout << "MultiGPU : Off" << std::endl;
  std::cout << rule << std::endl;
  std::cout << "OuterSolver : " << kUpper[outer] << std::endl;
  std::cout << "OuterMaxLoop : " << p.outer_maxloop << std::endl;
  std::cout << "OuterEps : " << p.outer_eps << std::endl;
  std::cout << "OuterRestart : " << p.outer_restart << std::endl;
  std::cout << "OuterKskip : " << p.outer_kskip << std::endl;
  std::cout << "OuterFix : " << p.outer_fix << std::endl;
  std::cout << rule << std::endl;
  std::cout << "InnerSolver : " << (inner ? kUpper[inner] : "NONE") << std::endl;
  std::cout << "InnerMaxLoop : " << p.inner_maxloop << std::endl;
  std::cout << "InnerEps : " << p.inner_eps << std::endl;
  std::cout << "InnerRestart : " << p.inner_restart << std::endl;
  std::cout << "InnerKskip : " << p.inner_kskip << std::endl;
  std::cout << "InnerFix : " << p.inner_fix << std::endl;
  std::cout << "=========================================================" << std::endl;

  ks_ctx* ctx = nullptr;
  int st = ks_ctx_create(std::atoi(get("device").c_str()), &ctx);
  if (st != KS_OK) { std::cerr << "[X] no GPU: " << ks_status_string(st) << std::endl; return 1; }
  uint32_t mflags = 0;
  const bool bicg_family = outer == 2 || outer == 8 || inner == 2 || inner == 8;
  if (bicg_family) mflags |= KS_MAT_TRANSPOSE;
  if (p.inner_precision == KS_F32) mflags |= KS_MAT_FP32_COPY;
  ks_mat* A = nullptr;
  st = ks_mat_create_csr(ctx, n, nnz, ptr.data(), col.data(), val.data(), mflags, &A);
  if (st != KS_OK) {
    std::cerr << "[X] matrix: " << ks_ctx_last_error(ctx) << std::endl;
    ks_ctx_destroy(ctx);
    return 1;
  }
  const int64_t hist_cap = (int64_t)std::max(p.outer_maxloop, 1) + 16;
  std::vector<double> hist(2 * hist_cap);
  ks_result res;
  std::fill(x.begin(), x.end(), 0.0);  // every reference solver zeroes x (cg.cpp:79)
  st = ks_solve(ctx, A, &p, b.data(), x.data(), hist.data(), hist_cap, &res);
  if (st != KS_OK) {
    std::cerr << "outer result -> " << st << " " << ks_ctx_last_error(ctx) << std::endl;
    std::cerr << "[X]Some Error in methods" << std::endl;
    ks_mat_destroy(A);
    ks_ctx_destroy(ctx);
    return 1;
  }
  const std::string name = kUpper[outer];
  FILE* fh = std::fopen(("./output/" + name + "_his.txt").c_str(), "w");
  FILE* fx = std::fopen(("./output/" + name + "_xvec.txt").c_str(), "w");
  if (!fh || !fx) { std::cerr << "File open error" << std::endl; return 1; }
  const int64_t nh = std::min(res.n_hist, hist_cap);
  for (int64_t i = 0; i < nh; ++i) {
    std::fprintf(fh, "%lld %.12E\n", (long long)hist[2 * i], hist[2 * i + 1]);
    if (p.verbose) std::printf("%lld %.12E\n", (long long)hist[2 * i], hist[2 * i + 1]);
  }
  for (int64_t i = 0; i < n; ++i) std::fprintf(fx, "%lld %.12E\n", (long long)i, x[i]);
  std::fclose(fh);
  std::fclose(fx);
  std::cout << "|b-ax|2/|b|2 = " << std::fixed << std::setprecision(1) << res.log10_check << std::endl;
  std::cout << "loop = " << res.iters << std::endl;
  std::cout << "time = " << std::setprecision(6) << res.seconds << std::endl;
  print_breakdown(res);
  if (res.status == KS_OK) std::cout << "converge" << std::endl;
  else if (res.status == KS_NOT_CONVERGED) std::cout << "not converge" << std::endl;
  else std::cerr << "outer result -> " << res.status << " (" << ks_status_string(res.status) << ")" << std::endl;
  if (bylong["json"]->set)
    std::printf("{\"solver\": \"%s\", \"status\": %d, \"iters\": %lld, \"relres_true\": %.6e, \"seconds\": %.6f}\n",
                name.c_str(), res.status, (long long)res.iters, res.relres_true, res.seconds);
  ks_mat_destroy(A);
  ks_ctx_destroy(ctx);
  return 0;
}
