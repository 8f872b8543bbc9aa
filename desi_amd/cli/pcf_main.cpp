// pcf_main.cpp -- the PCF command line, a drop-in for the reference's PCF.out
// (cuda/main.cu:45-429): same calculation types, flags, defaults, messages,
// input paths (relative to ../data/) and output files (in ../results/).
//
//   PCF <2iso|2ani|3iso|3ani> -n N -f datafile [-r randfile | -rd randdir/]
//       -b bins -d dmax [-bpc [-a]] [-p partitions] [-s size_box]
//
// The histograms are computed by libpcf_hip.so through include/pcf_abi.h.
// Differences from the reference, all fixes of defects:
//   * -rd lists the directory in sorted order (the reference uses readdir
//     order, cuda/main.cu:160-185);
//   * 3PCF output names do not accumulate across random files
//     (cuda/pcf3iso.cuh:127,132,137);
//   * the triplet weight is w1*w2*w3 per triplet (the reference multiplies
//     cumulatively, cuda/kernels/3iso_k.cuh:79);
//   * the 2ani DR weight uses the random point's weight
//     (cuda/kernels/2ani_k.cuh:128);
//   * 3ani writes the first bins^3 values of each histogram's own array (the
//     reference's writer reads a bins^3 stride into a bins^5 array for random
//     files after the first, cuda/pcf3ani.cuh:123-138); PCF_SAVE_5D=1 writes the
//     full bins^5 histogram with the reference's save_histogram5D format.
// --ref-bugs reproduces the reference's defects instead (for comparison with
// result files it wrote; pcf_set_ref_bugs for the library's part): the
// cumulative triplet weight and the XY2ani weight on the -p node grid, the
// 32-bit int overflows of the analytic normalisations, 3PCF output names that
// accumulate across random files (cuda/pcf3iso.cuh:127), -rd in readdir
// order (cuda/main.cu:164-185), and 3ani files of random file i holding values
// i*bins^3 .. (i+1)*bins^3 - 1 of the random files' concatenated bins^5
// histograms (save_histogram3D with nhistos = i, cuda/pcf3ani.cuh:123-138).
// Environment: PCF_DATA_DIR / PCF_RESULTS_DIR override ../data/ and
// ../results/; PCF_PRECISION=fp64 selects the plain FP64 path.
// PCF_TUNE_CACHE=<file> loads the pair kernels' tuning cache from the file at
// start (a missing file is an empty cache; one from another device model is
// reported and ignored) and writes it back at the end; with PCF_AUTOTUNE=1
// every pair histogram whose workload the cache does not hold is autotuned
// first (pcf_autotune), so the second run of a workload reuses the first's
// tuning.
// Extension: -e also writes the Landy-Szalay estimator of each random file,
// formed on the device by pcf_estimator as the reference's analysis scripts
// form it (cpp/2PCFisotropic/normal/graph_2DPF.py:8-14,
// cpp/3PCFisotropic/normal/graph_3PCF.py:5-6): LS<tag><random> for 2PCF,
// LS3<tag><random> for 3PCF (analytic modes: ...BPCanalytic_<data>), in the
// histogram's own file format; PCF_EST_NORMALIZE=1 divides the histograms by
// their pair (triplet) counts first.
#include <dirent.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "pcf_abi.h"
#include "pcf_io.h"

using std::cout;
using std::endl;
using std::string;

namespace {

void show_help() {
  cout << "Use: ./PCF <calc_type> -n number_of_points -f datafile -b bins -d maximum_distance"
       << endl;
  cout << "-r randomfile || -rd randoms_directory" << endl;
  cout << "[-bpc [-a]][-p box_partitions] [-s size_box]" << endl;
  cout << "Computes the histograms of the isotropic or anisotropic two or three point correlation\n"
          "function, with or without periodic boundary conditions, on an AMD Instinct GPU.\n"
          "Catalog files hold 4 columns: x, y, z and the weight of each point, one point per row,\n"
          "and must hold at least the number of points given with -n.  Give either -r or -rd.\n\n";
  cout << "Periodic boundaries: -bpc.  With -bpc -a (2iso and 3iso only) the random histograms are\n"
          "computed analytically and -r / -rd are not needed.  Without -s the box side is inferred\n"
          "as ceil(largest coordinate + 1), which may not be the true period.\n\n";
  cout << "<calc_type> options:\n\n"
          "3iso: Isotropic 3 points correlation function.\n"
          "3ani: Anisotropic 3 points correlation function.\n"
          "2iso: Isotropic 2 points correlation function.\n"
          "2ani: Anisotropic 2 points correlation function.\n";
  cout << "Parameters:\n\n"
          "-n number_of_points:   (required) points taken from each file.\n"
          "-f datafile:           (required) data file, relative to the data/ directory.\n"
          "-b bins:               (required) bins per histogram dimension.\n"
          "-d maximum_distance:   (required) largest separation counted.\n"
          "-r randomfile:         random file, relative to the data/ directory.\n"
          "-rd randoms_directory: directory (inside data/) of random files; the histograms that\n"
          "                       need randoms are computed for each file.\n"
          "-bpc:                  (optional) periodic boundary conditions.\n"
          "  [-a]:                (optional) analytic random histograms (isotropic types).\n"
          "-p box_partitions:     (optional) grid partitions per dimension; accepted for\n"
          "                       compatibility, the engine sizes its own grid.\n"
          "-s size_box:           (optional) side of the box holding the points.\n"
          "-e:                    (optional) also write the Landy-Szalay estimator per random\n"
          "                       file (LS... for 2PCF, LS3... for 3PCF).\n"
          "--ref-bugs:            (optional) reproduce the reference program's defects (weighted\n"
          "                       3PCF and 2ani DR weights, analytic overflows, file names), to\n"
          "                       compare with result files it wrote; -p then sets its node grid.\n\n";
}

struct Opts {
  string type;
  int np = 0, bn = 0, partitions = 35;
  bool bpc = false, analytic = false, rand_dir = false, size_box_provided = false;
  bool estimator = false;
  bool ref_bugs = false;
  float size_box = 0, dmax = 0;
  string data_name, rand_name;
};

string env_or(const char* k, const char* d) {
  const char* v = std::getenv(k);
  return v && *v ? string(v) : string(d);
}

string with_slash(string s) {
  if (!s.empty() && s.back() != '/') s += '/';
  return s;
}

int die(pcf_ctx* ctx, int rc, const string& what) {
  std::cerr << "PCF: " << what << " failed (" << rc << "): "
            << (ctx ? pcf_last_error(ctx) : pcf_create_error()) << endl;
  if (ctx) pcf_destroy(ctx);
  return rc == 0 ? 1 : (rc < 0 ? -rc + 1 : rc);
}

struct Run {
  pcf_ctx* ctx = nullptr;
  Opts o;
  string results;
  int precision = PCF_PREC_MIXED;

  pcf_job job(int type, int hist, pcf_catalog x, pcf_catalog y) const {
    pcf_job j;
    std::memset(&j, 0, sizeof(j));
    j.type = type;
    j.hist = hist;
    j.bpc = o.bpc ? 1 : 0;
    j.bins = o.bn;
    j.dmax = o.dmax;
    j.box = o.size_box;
    j.precision = precision;
    j.cat_x = x;
    j.cat_y = y;
    j.partitions = o.partitions;
    return j;
  }

  bool autotune = false;  // PCF_AUTOTUNE=1

  int hist(int type, int h, pcf_catalog x, pcf_catalog y, std::vector<double>* out) {
    int64_t n = pcf_hist_size(type, o.bn);
    out->assign((size_t)n, 0.0);
    pcf_job j = job(type, h, x, y);
    if (autotune && (type == PCF_2ISO || type == PCF_2ANI) && precision == PCF_PREC_MIXED) {
      const int found = pcf_tune_lookup(ctx, &j, 1, nullptr);
      if (found < 0) return found;
      if (found == 0) {
        const int r = pcf_autotune(ctx, &j, 1, nullptr);
        if (r != PCF_OK) return r;
      }
    }
    return pcf_histogram(ctx, &j, nullptr, out->data());
  }

  // Landy-Szalay estimator of finished histograms (2PCF: DD, DR, RR; 3PCF:
  // DDD, DDR, DRR, RRR), written like a histogram of this type.
  int estimate(int type, const std::vector<const std::vector<double>*>& hs, const string& name) {
    const int order = (type == PCF_2ISO || type == PCF_2ANI) ? 2 : 3;
    const size_t n = hs[0]->size();
    std::vector<double> xi(n);
    const int flags = env_or("PCF_EST_NORMALIZE", "0") == "1" ? PCF_EST_NORMALIZE : 0;
    int r = pcf_estimator(ctx, order, (int64_t)n, hs[0]->data(), hs[1]->data(), hs[2]->data(),
                          order == 3 ? hs[3]->data() : nullptr, (double)o.np, (double)o.np, flags,
                          xi.data());
    if (r != PCF_OK) return r;
    return save(type, name, xi) ? PCF_OK : PCF_ERR_ARG;
  }

  bool save(int type, const string& name, const std::vector<double>& h) {
    string path = results + name;
    switch (type) {
      case PCF_2ISO: return pcfcli::save_histogram1D(path, o.bn, h.data());
      case PCF_2ANI: return pcfcli::save_histogram2D(path, o.bn, h.data());
      case PCF_3ISO: return pcfcli::save_histogram3D(path, o.bn, h.data());
      case PCF_3ANI:
        if (env_or("PCF_SAVE_5D", "0") == "1")
          return pcfcli::save_histogram5D(path, o.bn, h.data());
        return pcfcli::save_histogram3D(path, o.bn, h.data());
    }
    return false;
  }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc == 2 && (std::strcmp(argv[1], "--help") == 0 || std::strcmp(argv[1], "-h") == 0)) {
    show_help();
    return 0;
  }
  const bool type_ok = argc >= 8 && (std::strcmp(argv[1], "3iso") == 0 ||
                                     std::strcmp(argv[1], "3ani") == 0 ||
                                     std::strcmp(argv[1], "2iso") == 0 ||
                                     std::strcmp(argv[1], "2ani") == 0);
  if (!type_ok) {
    cout << "Invalid <calc_type> option or not enough parameters. \nSee --help for more information."
         << endl;
    return 1;
  }
  Opts o;
  o.type = argv[1];
  try {
    for (int i = 2; i < argc; i++) {
      string a = argv[i];
      auto next = [&]() -> string {
        if (i + 1 >= argc) throw std::invalid_argument(a);
        return string(argv[++i]);
      };
      if (a == "-n") o.np = std::stoi(next());
      else if (a == "-f") o.data_name = next();
      else if (a == "-r" || a == "-rd") {
        o.rand_dir = (a == "-rd");
        o.rand_name = next();
      } else if (a == "-b") o.bn = std::stoi(next());
      else if (a == "-d") o.dmax = std::stof(next());
      else if (a == "-bpc") o.bpc = true;
      else if (a == "-a") o.analytic = true;
      else if (a == "-e") o.estimator = true;
      else if (a == "--ref-bugs") o.ref_bugs = true;
      else if (a == "-p") o.partitions = (int)std::stof(next());
      else if (a == "-s") {
        o.size_box = std::stof(next());
        o.size_box_provided = true;
      }
    }
  } catch (const std::exception&) {
    cout << "Invalid value for a command line parameter." << endl;
    return 1;
  }
  // Required arguments (cuda/main.cu:99-117).
  if (o.np == 0) {
    cout << "Missing number of points argument." << endl;
    return 1;
  }
  if (o.bn == 0) {
    cout << "Missing number of bins argument." << endl;
    return 1;
  }
  if (o.dmax == 0) {
    cout << "Missing maximum distance argument." << endl;
    return 1;
  }
  const bool iso = (o.type == "3iso" || o.type == "2iso");
  const bool rand_required = !(o.bpc && o.analytic && iso);
  if (rand_required && o.rand_name.empty()) {
    cout << "Missing random file(s) location." << endl;
    return 1;
  }

  const string data_dir = with_slash(env_or("PCF_DATA_DIR", "../data/"));
  Run run;
  run.o = o;
  run.results = with_slash(env_or("PCF_RESULTS_DIR", "../results/"));
  run.precision = env_or("PCF_PRECISION", "mixed") == "fp64" ? PCF_PREC_FP64 : PCF_PREC_MIXED;
  auto t_host0 = std::chrono::steady_clock::now();

  // Read the catalogs (cuda/main.cu:148-213).
  string err;
  std::vector<float> D;
  double maxc = 0;
  if (!pcfcli::read_catalog(data_dir + o.data_name, o.np, &D, &maxc, &err)) {
    cout << err << endl;
    return 1;
  }
  if (run.o.size_box <= 0) run.o.size_box = (float)std::ceil(maxc + 1);
  std::vector<string> rand_files, histo_names;
  histo_names.push_back(o.data_name);
  if (rand_required) {
    if (o.rand_dir) {
      string dir = data_dir + o.rand_name;
      DIR* folder = opendir(dir.c_str());
      if (!folder) {
        cout << "Failed to open the randoms directory " << dir << endl;
        return 1;
      }
      std::vector<string> names;
      while (dirent* ent = readdir(folder)) {
        string nm = ent->d_name;
        if (nm != "." && nm != "..") names.push_back(nm);
      }
      closedir(folder);
      if (!o.ref_bugs) std::sort(names.begin(), names.end());
      for (auto& nm : names) {
        histo_names.push_back(nm);
        rand_files.push_back(o.rand_name + nm);
      }
    } else {
      rand_files.push_back(o.rand_name);
      histo_names.push_back(o.rand_name);
    }
  }
  std::vector<std::vector<float>> R(rand_files.size());
  float r_size_box = 0;
  for (size_t i = 0; i < rand_files.size(); ++i) {
    double rmax = 0;
    if (!pcfcli::read_catalog(data_dir + rand_files[i], o.np, &R[i], &rmax, &err)) {
      cout << err << endl;
      return 1;
    }
    // The reference keeps the first random file's inferred box (its
    // open_files only sets a box that is still <= 0, cuda/create_grid.cuh:75).
    if (r_size_box <= 0) r_size_box = (float)std::ceil(rmax + 1);
    if (!o.size_box_provided && r_size_box > run.o.size_box) run.o.size_box = r_size_box;
  }

  int rc = pcf_create(&run.ctx, 0);
  if (rc != PCF_OK) return die(nullptr, rc, "pcf_create");
  if (o.ref_bugs && (rc = pcf_set_ref_bugs(run.ctx, 1)) != PCF_OK)
    return die(run.ctx, rc, "pcf_set_ref_bugs");
  const string tune_cache = env_or("PCF_TUNE_CACHE", "");
  run.autotune = env_or("PCF_AUTOTUNE", "0") == "1";
  if (!tune_cache.empty()) {
    FILE* tf = std::fopen(tune_cache.c_str(), "r");
    if (tf) {
      std::fclose(tf);
      if (pcf_tune_load(run.ctx, tune_cache.c_str()) != PCF_OK)
        std::cerr << "PCF: ignoring the tuning cache: " << pcf_last_error(run.ctx) << endl;
    }
  }
  pcf_catalog cd = -1;
  std::vector<pcf_catalog> cr(R.size(), -1);
  if ((rc = pcf_catalog_upload(run.ctx, D.data(), o.np, &cd)) != PCF_OK)
    return die(run.ctx, rc, "data upload");
  for (size_t i = 0; i < R.size(); ++i)
    if ((rc = pcf_catalog_upload(run.ctx, R[i].data(), o.np, &cr[i])) != PCF_OK)
      return die(run.ctx, rc, "random upload");
  auto t_host1 = std::chrono::steady_clock::now();
  cout << "Succesfully readed the data. All set to compute the histograms in "
       << std::chrono::duration<double, std::milli>(t_host1 - t_host0).count() << " miliseconds"
       << endl;

  const string& dn = histo_names[0];
  std::vector<double> h;
  auto t0 = std::chrono::steady_clock::now();
  auto H = [&](int type, int hist, pcf_catalog x, pcf_catalog y, const string& name,
               std::vector<double>* keep = nullptr) -> int {
    int r = run.hist(type, hist, x, y, &h);
    if (r != PCF_OK) return r;
    if (keep) *keep = h;
    return run.save(type, name, h) ? PCF_OK : PCF_ERR_ARG;
  };
  std::vector<double> h0, h1, h2, h3;  // kept for the estimator (-e)

  if (o.type == "2iso" || o.type == "2ani") {
    const int type = (o.type == "2iso") ? PCF_2ISO : PCF_2ANI;
    const string tag = (type == PCF_2ISO ? "iso_" : "ani_") + string(o.bpc ? "BPC_" : "");
    if (type == PCF_2ISO && o.bpc && o.analytic) {
      // cuda/pcf2isoBPC.cuh:166-289: DD with periodic images, RR analytic, DR = RR.
      if ((rc = H(type, PCF_XX, cd, -1, "DDiso_BPCanalytic_" + dn, &h0)) != PCF_OK)
        return die(run.ctx, rc, "DD");
      std::vector<double> rr(o.bn);
      if ((rc = pcf_analytic_rr(run.ctx, o.bn, o.np, o.dmax, run.o.size_box, rr.data())) !=
          PCF_OK)
        return die(run.ctx, rc, "analytic RR");
      run.save(type, "RRiso_BPCanalytic_" + dn, rr);
      run.save(type, "DRiso_BPCanalytic_" + dn, rr);
      if (o.estimator && (rc = run.estimate(type, {&h0, &rr, &rr}, "LSiso_BPCanalytic_" + dn)) !=
                             PCF_OK)
        return die(run.ctx, rc, "estimator");
    } else {
      if ((rc = H(type, PCF_XX, cd, -1, "DD" + tag + dn, &h0)) != PCF_OK)
        return die(run.ctx, rc, "DD");
      for (size_t i = 0; i < cr.size(); ++i) {
        if ((rc = H(type, PCF_XX, cr[i], -1, "RR" + tag + histo_names[i + 1], &h2)) != PCF_OK)
          return die(run.ctx, rc, "RR");
        if ((rc = H(type, PCF_XY, cd, cr[i], "DR" + tag + histo_names[i + 1], &h1)) != PCF_OK)
          return die(run.ctx, rc, "DR");
        if (o.estimator &&
            (rc = run.estimate(type, {&h0, &h1, &h2}, "LS" + tag + histo_names[i + 1])) != PCF_OK)
          return die(run.ctx, rc, "estimator");
      }
    }
  } else {
    const int type = (o.type == "3iso") ? PCF_3ISO : PCF_3ANI;
    const string tag = (type == PCF_3ISO ? "iso_" : "ani_") + string(o.bpc ? "BPC_" : "");
    if (type == PCF_3ISO && o.bpc && o.analytic) {
      // cuda/pcf3isoBPC.cuh:177-375: periodic DDD, analytic RRR and DDR, DRR = DDR.
      if ((rc = H(type, PCF_XXX, cd, -1, "DDDiso_BPCanalytic_" + dn, &h0)) != PCF_OK)
        return die(run.ctx, rc, "DDD");
      const size_t b3 = (size_t)o.bn * o.bn * o.bn;
      std::vector<double> ddr(b3), rrr(b3);
      if ((rc = pcf_analytic_3iso(run.ctx, cd, o.bn, o.dmax, run.o.size_box, o.np, 0, 0,
                                  ddr.data(), rrr.data(), nullptr, nullptr)) != PCF_OK)
        return die(run.ctx, rc, "analytic DDR/RRR");
      run.save(type, "DDRiso_BPCanalytic_" + dn, ddr);
      run.save(type, "DRRiso_BPCanalytic_" + dn, ddr);
      run.save(type, "RRRiso_BPCanalytic_" + dn, rrr);
      if (o.estimator &&
          (rc = run.estimate(type, {&h0, &ddr, &ddr, &rrr}, "LS3iso_BPCanalytic_" + dn)) != PCF_OK)
        return die(run.ctx, rc, "estimator");
      auto t1 = std::chrono::steady_clock::now();
      cout << "Spent " << std::chrono::duration<double, std::milli>(t1 - t0).count()
           << " miliseconds to compute and save all the histograms." << endl;
      pcf_destroy(run.ctx);
      cout << "Program terminated..." << endl;
      return 0;
    }
    if ((rc = H(type, PCF_XXX, cd, -1, "DDD" + tag + dn, &h0)) != PCF_OK)
      return die(run.ctx, rc, "DDD");
    // --ref-bugs: the names accumulate the random files' names, and 3ani
    // random file i writes the slice i*bins^3 of the concatenated histograms.
    string acc_rrr = "RRR" + tag, acc_ddr = "DDR" + tag, acc_drr = "DRR" + tag;
    std::vector<std::vector<double>> all_rrr, all_ddr, all_drr;
    const size_t b3 = (size_t)o.bn * o.bn * o.bn;
    auto slab = [&](const std::vector<std::vector<double>>& all, size_t i) {
      const size_t off = i * b3, per = all[0].size();
      std::vector<double> v(b3, 0.0);
      for (size_t k = 0; k < b3; ++k) {
        const size_t g = off + k;
        if (g / per < all.size()) v[k] = all[g / per][g % per];
      }
      return v;
    };
    auto H3 = [&](int hist, pcf_catalog x, pcf_catalog y, const string& name,
                  std::vector<double>* keep, std::vector<std::vector<double>>* all,
                  size_t i) -> int {
      if (!(o.ref_bugs && type == PCF_3ANI)) return H(type, hist, x, y, name, keep);
      int r = run.hist(type, hist, x, y, &h);
      if (r != PCF_OK) return r;
      *keep = h;
      all->push_back(h);
      // Random file i's slice lies in a histogram already computed (i*bins^3
      // < (i+1)*bins^5).
      return run.save(PCF_3ISO, name, slab(*all, i)) ? PCF_OK : PCF_ERR_ARG;
    };
    for (size_t i = 0; i < cr.size(); ++i) {
      const string& rn = histo_names[i + 1];
      acc_rrr += rn;
      acc_ddr += rn;
      acc_drr += rn;
      const string n_rrr = o.ref_bugs ? acc_rrr : "RRR" + tag + rn;
      const string n_ddr = o.ref_bugs ? acc_ddr : "DDR" + tag + rn;
      const string n_drr = o.ref_bugs ? acc_drr : "DRR" + tag + rn;
      if ((rc = H3(PCF_XXX, cr[i], -1, n_rrr, &h3, &all_rrr, i)) != PCF_OK)
        return die(run.ctx, rc, "RRR");
      if ((rc = H3(PCF_XXY, cd, cr[i], n_ddr, &h1, &all_ddr, i)) != PCF_OK)
        return die(run.ctx, rc, "DDR");
      if ((rc = H3(PCF_XXY, cr[i], cd, n_drr, &h2, &all_drr, i)) != PCF_OK)
        return die(run.ctx, rc, "DRR");
      if (o.estimator && (rc = run.estimate(type, {&h0, &h1, &h2, &h3}, "LS3" + tag + rn)) !=
                             PCF_OK)
        return die(run.ctx, rc, "estimator");
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  cout << "Spent " << std::chrono::duration<double, std::milli>(t1 - t0).count()
       << " miliseconds to compute and save all the histograms." << endl;
  if (!tune_cache.empty() && pcf_tune_save(run.ctx, tune_cache.c_str()) != PCF_OK)
    std::cerr << "PCF: " << pcf_last_error(run.ctx) << endl;
  pcf_destroy(run.ctx);
  cout << "Program terminated..." << endl;
  return 0;
}
